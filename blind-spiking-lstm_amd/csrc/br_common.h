// Pieces shared by the FP64 blind-rotation kernels (blind_rotate.hip: one
// workgroup per ciphertext or pair of ciphertexts; br_pair.hip: a pair of
// workgroups per ciphertext): key-slot selection, the key loads of one bin
// pair, the multiply-accumulate of one pattern, and the accumulator modes.
#pragma once

#include "fold_fft.h"

namespace bslstm {

// Per-ciphertext key selection for batches that mix clients: the key of
// ciphertext c is bsk + slots[c] * stride (complex elements), slots clamped
// to [0, count).  slots == nullptr: one key for the whole batch.  With two
// ciphertexts per workgroup both must use the same slot (the slot of the
// first is used).
struct KeySlots {
    const int32_t *slots;
    int64_t stride;
    int32_t count;
};

// BSK values of one bin pair for pattern p of gadget level lev, output rows
// a0 .. a0+NA-1: [a][bc2][0/1] = F-bins 2r, 2r+1 of BSK[g, a0+a, 2*lev+bc2, p].
// The block pointer is uniform across the workgroup and the lane offset (2r)
// is 32-bit.
template <int L, int NA>
__device__ __forceinline__ void load_pair(cplx (&dst)[NA][2][2], const cplx *__restrict__ bsk_g,
                                          int r2, int a0, int lev, int p, int m, int M)
{
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b2 = 0; b2 < 2; b2++) {
            const cplx *blk = bsk_g + ((size_t)((a0 + a) * 2 * L + 2 * lev + b2) * m + p) * M;
            dst[a][b2][0] = blk[r2];
            dst[a][b2][1] = blk[r2 + 1];
        }
}

// W[a][bc2][c][.] += BSK(p) * rotation root of ciphertext c, for bins r and
// M-1-r (the latter in conjugate form, see tfhe_device.h).  Each root serves
// the NA output rows held.
template <int N, int CT, int NA, bool FULL>
__device__ __forceinline__ void mac_pattern(cplx (&W)[NA][2][CT][2], const cplx (&bk)[NA][2][2],
                                            const int32_t *const (&idp)[CT], int p, int r,
                                            const cplx *ztab)
{
#pragma unroll
    for (int c = 0; c < CT; c++) {
        // exponents matter mod 2N: a 24-bit multiply (operands < 2^13)
        const int e = (int)((uint32_t)idp[c][p] & (2 * N - 1));
        const int e0 = (int)__umul24((uint32_t)e, (uint32_t)(4 * r + 1));
        const cplx rot0 = zeta_pow<N, FULL>(ztab, e0);
        // zeta^(e(4(M-1-r)+1)) = zeta^(-e(4r+3)) = conj(zeta^(e(4r+1)) * zeta^(2e)),
        // zeta^(2e) being uniform across the workgroup (one broadcast read)
        const cplx rot1 = cconj(cmul(rot0, zeta_pow<N, FULL>(ztab, 2 * e)));
#pragma unroll
        for (int a = 0; a < NA; a++)
#pragma unroll
            for (int b2 = 0; b2 < 2; b2++) {
                W[a][b2][c][0] = cfma(bk[a][b2][0], rot0, W[a][b2][c][0]);
                W[a][b2][c][1] = cfma_conj(bk[a][b2][1], rot1, W[a][b2][c][1]);
            }
    }
}

// ACC: how the accumulator sits in registers.
//  ACC_32  log2_q <= 32: its class mod 2^32 in 32-bit registers.
//  ACC_HI  log2_q > 32 with l*log2_beta + 1 <= 32: bits [log2_q-32, log2_q) of
//          the class in 32-bit registers.  The gadget digits read only the top
//          l*log2_beta bits plus the one rounding bit below them (every offset
//          term is a multiple of 2^(log2_q-32)), so those bits are all a step
//          consumes; the full class is written out at the last step.
//  ACC_64  otherwise: all 64 bits.
enum { ACC_64 = 0, ACC_32 = 1, ACC_HI = 2 };

}  // namespace bslstm
