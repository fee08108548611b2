// Status codes and per-thread error text of the C ABI.
#include <cstdarg>
#include <cstdio>

#include "ksm_common.h"

namespace {
thread_local char g_last_error[512] = "";
}

namespace ksm {

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *where) {
    return fail(KSM_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

}  // namespace ksm

extern "C" {

int ksm_version(void) { return 100; }

const char *ksm_status_string(int status) {
    switch (status) {
        case KSM_OK: return "ok";
        case KSM_ERR_INVALID: return "invalid argument";
        case KSM_ERR_CAPACITY: return "spike capacity exceeded";
        case KSM_ERR_HIP: return "HIP runtime error";
        case KSM_ERR_WORKSPACE: return "workspace too small";
        case KSM_ERR_TIES: return "exact score ties exceed slot capacity";
        default: return "unknown status";
    }
}

const char *ksm_last_error(void) { return g_last_error; }

}  // extern "C"
