// Library state, error reporting and device selection of the C ABI.
#include "common.hpp"

namespace dwba {

namespace {
thread_local std::string g_err;
int g_init = 0;
int g_cu = 256;
}  // namespace

std::string& last_error() { return g_err; }

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

bool initialised() { return g_init != 0; }
int device_cu_count() { return g_cu; }

}  // namespace dwba

extern "C" {

int32_t dwba_version(void) { return 100; }

int32_t dwba_init(int32_t device) {
    int n = 0;
    DWBA_HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return dwba::set_error(DWBA_ERR_ARG, "dwba_init: device %d out of range (%d devices)", device, n);
    DWBA_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t p;
    DWBA_HIP_CHECK(hipGetDeviceProperties(&p, device));
    dwba::g_cu = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
    const int rc = dwba::noalloc_scratch_setup(device);
    if (rc != DWBA_OK) return rc;
    dwba::g_init = 1;
    dwba::g_err.clear();
    return DWBA_OK;
}

int32_t dwba_shutdown(void) {
    dwba::g_init = 0;
    return DWBA_OK;
}

const char* dwba_last_error(void) { return dwba::g_err.c_str(); }

}  // extern "C"
