// Host-side helpers shared by the epcq HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/epcq_abi.h"

namespace epcq {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

// Turns a pending launch error into EPCQ_ERR_HIP with the call's name.
int check_launch(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int ceil_log2(int64_t x) {
    int b = 0;
    while ((int64_t(1) << b) < x) ++b;
    return b;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace epcq
