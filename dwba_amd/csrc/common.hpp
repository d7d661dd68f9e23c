// Shared host/device helpers of the gfx950 radial-integral library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/dwba_hip.h"

namespace dwba {

// Thread-local message of the last failing call (returned by dwba_last_error).
std::string& last_error();
int set_error(int code, const char* fmt, ...);
bool initialised();

#define DWBA_HIP_CHECK(expr)                                                        \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::dwba::set_error(DWBA_ERR_HIP, "%s failed: %s (%s:%d)", #expr,  \
                                     hipGetErrorString(e_), __FILE__, __LINE__);    \
    } while (0)

#define DWBA_LAUNCH_CHECK(name)                                                     \
    do {                                                                            \
        hipError_t e_ = hipGetLastError();                                          \
        if (e_ != hipSuccess)                                                       \
            return ::dwba::set_error(DWBA_ERR_HIP, "launch of %s failed: %s", name, \
                                     hipGetErrorString(e_));                        \
    } while (0)

#define DWBA_REQUIRE(cond, ...)                                                     \
    do {                                                                            \
        if (!(cond)) return ::dwba::set_error(DWBA_ERR_ARG, __VA_ARGS__);           \
    } while (0)

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// Number of CUs of the current device (recorded by dwba_init).
int device_cu_count();

// Records the current device's address of the fixed scratch array of
// dwba_il_plain_f64_noalloc (plain_f64.hip); called by dwba_init.
int noalloc_scratch_setup(int device);

// Deterministic wavefront-wide sum of a double (xor butterfly over 64 lanes).
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

}  // namespace dwba
