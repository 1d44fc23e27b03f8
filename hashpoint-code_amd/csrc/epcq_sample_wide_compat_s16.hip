// Sampling kernels for windows of 65..256 lanes (k <= 16), neighbour mode compat, sample bound 16 (see epcq_sample_kernels.h).
#include "epcq_sample_kernels.h"

namespace epcq {
EPCQ_INSTANTIATE_LAUNCH(EPCQ_MODE_COMPAT, kCompatBuffer, kRegSamples, 4)
}  // namespace epcq
