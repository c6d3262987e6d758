// Device helpers shared by the gfx950 TFHE kernels: complex FP64 arithmetic
// and the quarter-wave root table that serves every root of unity the
// negacyclic transform needs.
//
// Transform convention (the reference's, csrc/dec_fft.cuh:33-57 and
// schemas/polynomial_jax.py:11-21): for a real polynomial p of degree < N,
//     F_j = p(zeta^(2j+1)),  zeta = exp(-i*pi/N),  j < N/2.
// The kernels work in the folded domain
//     U_r = p(zeta^(4r+1)),  r < N/2,
// computed as an N/2-point complex DFT of u_k = (p_k - i*p_{k+N/2}) * zeta^k.
// The two domains hold the same values: U_r = F_{2r} for r < N/4 and
// U_r = conj(F_{N-1-2r}) for r >= N/4, so the U-bins r and N/2-1-r are the
// F-bins 2r and 2r+1 -- adjacent in memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bslstm {

struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return {fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
// acc + a*b
__device__ __forceinline__ cplx cfma(cplx a, cplx b, cplx acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
    return acc;
}
// acc + conj(a)*b
__device__ __forceinline__ cplx cfma_conj(cplx a, cplx b, cplx acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
    return acc;
}
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }
// multiply by -i / +i
__device__ __forceinline__ cplx cmul_mi(cplx a) { return {a.im, -a.re}; }
__device__ __forceinline__ cplx cmul_pi(cplx a) { return {-a.im, a.re}; }

// Root table in LDS: zeta^r = exp(-i*pi*r/N) for r in [0, N), stored at
// root_slot(r).  The slot XORs the low four index bits with a fold of the
// higher ones: lanes reading roots whose exponents step by a power of two
// (rotation roots e*(4r+1), transform twiddles) then spread over the LDS
// bank groups instead of piling onto one.
template <int N>
__device__ __forceinline__ int root_slot(int r)
{
    int h = r >> 4, f = 0;
#pragma unroll
    for (int b = 4; b < 13; b += 4) {
        f ^= h & 15;
        h >>= 4;
    }
    return r ^ f;
}

// FULL: the table holds the whole period, zeta^r for r in [0, 2N) (2N
// entries, twice the LDS), so a lookup needs no sign flip.
template <int N, bool FULL = false>
struct RootTab {
    static constexpr int SIZE = FULL ? 2 * N : N;
};

template <int N, bool FULL = false>
__device__ __forceinline__ void fill_root_table(cplx *ztab)
{
    for (int r = threadIdx.x; r < RootTab<N, FULL>::SIZE; r += blockDim.x) {
        double s, c;
        sincospi((double)r / (double)N, &s, &c);
        ztab[root_slot<N>(r)] = {c, -s};
    }
}

// zeta^e for any integer e.  Half-period table: zeta^N = -1, so zeta^e =
// +-zeta^(e mod N), the sign given by bit log2(N) of e -- one 16-byte LDS
// read and a sign flip of both halves (xor on the high words).  Full-period
// table: one read of zeta^(e mod 2N).
template <int N, bool FULL = false>
__device__ __forceinline__ cplx zeta_pow(const cplx *ztab, int e)
{
    if constexpr (FULL) {
        return ztab[root_slot<N>(e & (2 * N - 1))];
    } else {
        const cplx z = ztab[root_slot<N>(e & (N - 1))];
        const uint32_t flip = ((uint32_t)e & N) ? 0x80000000u : 0u;
        return {__hiloint2double((int)(__double2hiint(z.re) ^ flip), __double2loint(z.re)),
                __hiloint2double((int)(__double2hiint(z.im) ^ flip), __double2loint(z.im))};
    }
}

}  // namespace bslstm
