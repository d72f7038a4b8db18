// ABI housekeeping: version, thread-local error messages, device probe.
#include <cstdio>
#include <cstring>

#include "common.h"

namespace etm {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
}

}  // namespace etm

extern "C" int etm_abi_version(void) { return ETM_ABI_VERSION; }

extern "C" int etm_last_error(char* h_buf, size_t n) {
    ETM_ARG(h_buf != nullptr && n > 0, "etm_last_error: empty buffer");
    strncpy(h_buf, etm::g_last_error, n - 1);
    h_buf[n - 1] = '\0';
    return ETM_OK;
}

extern "C" int etm_device_count(int* h_out) {
    ETM_ARG(h_out != nullptr, "etm_device_count: null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *h_out = n;
    return ETM_OK;
}
