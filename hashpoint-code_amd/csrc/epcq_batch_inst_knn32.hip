// Persistent batched sampling kernels of neighbour mode knn32 (see epcq_batch_kernels.h).
#include "epcq_batch_kernels.h"

EPCQ_INSTANTIATE_BATCH(EPCQ_MODE_KNN, 32)
