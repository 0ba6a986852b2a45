// Library-wide C ABI: version, per-thread error text, device query.
#include <stdarg.h>

#include "dl_common.hpp"

namespace dl {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace dl

extern "C" {

int32_t dl_abi_version(void) { return DL_ABI_VERSION; }

int32_t dl_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return DL_E_ARG;
    snprintf(buf, len, "%s", dl::g_err);
    return 0;
}

int32_t dl_device_count(int32_t *count) {
    DL_CHECK_ARG(count != nullptr, "dl_device_count: NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return 0;
}

}  // extern "C"
