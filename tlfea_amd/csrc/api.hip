// General C-ABI entry points: version, error reporting, sizes.
#include <cstdarg>

#include "common.h"

namespace tlfea_hip {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace tlfea_hip

extern "C" const char* tlfea_version(void) { return "tlfea-hip 0.1 (gfx950)"; }

extern "C" const char* tlfea_last_error(void) { return tlfea_hip::g_err; }

extern "C" int tlfea_device_synchronize(void) {
    TLFEA_HIP_TRY(hipDeviceSynchronize());
    return TLFEA_OK;
}

extern "C" int64_t tlfea_padded_elems(int n_elem) { return tlfea_hip::padded(n_elem < 0 ? 0 : n_elem); }
