// Error state of the C ABI (include/bslstm_tfhe.h).
#include <stdarg.h>
#include <stdio.h>

#include "capi_internal.h"

namespace bslstm {

static thread_local char g_last_error[512];

int fail(int status, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return status;
}

void clear_error() { g_last_error[0] = '\0'; }

}  // namespace bslstm

extern "C" const char *bslstm_last_error(void) { return bslstm::g_last_error; }

extern "C" int bslstm_abi_version(void) { return 2; }
