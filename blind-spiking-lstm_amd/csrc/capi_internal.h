// Host-side helpers shared by the C-ABI translation units: per-thread last
// error message and status shorthands.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bslstm_tfhe.h"

namespace bslstm {

int fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

inline int launch_status(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(BSLSTM_INTERNAL, "%s launch failed: %s", what, hipGetErrorString(e));
    return BSLSTM_OK;
}

inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace bslstm
