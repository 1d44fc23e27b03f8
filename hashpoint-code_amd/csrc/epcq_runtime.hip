// Error reporting and small utilities of the epcq C ABI.
#include <string>

#include "epcq_common.h"
#include "epcq_device.h"

namespace epcq {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return EPCQ_OK;
    return fail(EPCQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
__global__ void debug_expf_kernel(const float* __restrict__ in, float* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = expf_cw(in[i]);
}
}  // namespace

}  // namespace epcq

extern "C" int epcq_abi_version(void) { return EPCQ_ABI_VERSION; }

extern "C" const char* epcq_last_error(void) { return epcq::g_last_error.c_str(); }

extern "C" int epcq_debug_expf(const float* in, float* out, int64_t n, void* stream) {
    if (n < 0) return epcq::fail(EPCQ_ERR_INVALID, "debug_expf: negative size");
    if (n == 0) return EPCQ_OK;
    const unsigned blocks = (unsigned)epcq::ceil_div(n, 256);
    epcq::debug_expf_kernel<<<dim3(blocks), dim3(256), 0, epcq::as_stream(stream)>>>(in, out, n);
    return epcq::check_launch("debug_expf");
}
