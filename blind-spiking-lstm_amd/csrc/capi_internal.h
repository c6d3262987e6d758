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

// Host-side validation shared by the blind-rotation entry points
// (blind_rotate.hip; mul_bootstrap.cuh:138-228 plus the shape checks the
// reference does not make).  *stride = 0 (broadcast test polynomial) or 2N.
int validate_blind_rotation(int64_t ct_elems, const double *bsk, const int64_t *bsk_dims,
                            const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                            const int32_t *indices, int64_t idx_elems, int64_t *output,
                            int64_t batch, int64_t log2_beta, int64_t log2_q, int64_t l,
                            int64_t degree, const int64_t *ct_init, int64_t *stride);

}  // namespace bslstm
