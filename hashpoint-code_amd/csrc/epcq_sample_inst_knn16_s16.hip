// Sampling kernels of neighbour mode knn16, sample bound 16 (see epcq_sample_kernels.h).
#include "epcq_sample_kernels.h"

namespace epcq {
EPCQ_INSTANTIATE_LAUNCH(EPCQ_MODE_KNN, 16, kRegSamples, 1)
}  // namespace epcq
