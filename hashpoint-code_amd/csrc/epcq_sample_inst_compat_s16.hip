// Sampling kernels of neighbour mode compat, sample bound 16 (see epcq_sample_kernels.h).
#include "epcq_sample_kernels.h"

namespace epcq {
EPCQ_INSTANTIATE_LAUNCH(EPCQ_MODE_COMPAT, kCompatBuffer, kRegSamples, 1)
}  // namespace epcq
