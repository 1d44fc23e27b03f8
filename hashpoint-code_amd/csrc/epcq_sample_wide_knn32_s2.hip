// Sampling kernels for windows of 65..256 lanes (k <= 16), neighbour mode knn32, sample bound 2 (see epcq_sample_kernels.h).
#include "epcq_sample_kernels.h"

namespace epcq {
EPCQ_INSTANTIATE_LAUNCH(EPCQ_MODE_KNN, 32, 2, 4)
}  // namespace epcq
