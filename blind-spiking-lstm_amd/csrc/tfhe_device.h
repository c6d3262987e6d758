// Device helpers shared by the gfx950 TFHE kernels: complex FP64 arithmetic,
// the quarter-wave cosine table that serves every root of unity the
// negacyclic transform needs, and the folded N/2-point negacyclic transform
// conventions.
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

// Quarter-wave table ctab[x] = cos(pi*x/N), x in [0, N/2], in LDS.
template <int N>
__device__ __forceinline__ double cos_pi_over_n(const double *ctab, int e)
{
    // e in [0, 2N)
    int e1 = e <= N ? e : 2 * N - e;                 // cos is even about pi
    double v = ctab[e1 <= N / 2 ? e1 : N - e1];
    return e1 <= N / 2 ? v : -v;
}

// zeta^e = exp(-i*pi*e/N) for any integer e.
template <int N>
__device__ __forceinline__ cplx zeta_pow(const double *ctab, int e)
{
    e &= 2 * N - 1;
    double c = cos_pi_over_n<N>(ctab, e);
    double s = cos_pi_over_n<N>(ctab, (e - N / 2) & (2 * N - 1));   // sin(pi e/N)
    return {c, -s};
}

template <int N>
__device__ __forceinline__ void fill_cos_table(double *ctab)
{
    for (int x = threadIdx.x; x <= N / 2; x += blockDim.x)
        ctab[x] = cospi((double)x / (double)N);
}

}  // namespace bslstm
