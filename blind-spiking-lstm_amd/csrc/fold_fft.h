// Folded negacyclic FP64 transform for gfx950 workgroups: one real
// length-N polynomial is one N/2-point complex FFT of
// u_k = (p_k - i p_{k+N/2}) zeta^k (the reference's F_j = p(zeta^(2j+1))
// convention, csrc/dec_fft.cuh:33-57, up to the bin order given in
// tfhe_device.h).  A group of T = N/8 threads transforms NP (one or two)
// polynomials at once, stored M apart; thread t holds points t + r*T (r < 4)
// of each in registers.
//
//  * forward_pair / inverse_pair: Stockham radix-4 (+ one radix-2 pass when
//    log2(N/2) is odd) through LDS, natural bin order, any N in 128..4096.
//  * forward_wave / inverse_wave (N in 512..2048): one radix-W cross-wave
//    stage through LDS, then four in-wave radix-4 passes over 256-point
//    regions whose data move between passes by cross-lane exchanges
//    (v_permlane32/16_swap, DPP), not LDS; bin b lives at phys(b).
#pragma once

#include "tfhe_device.h"

namespace bslstm {

template <int N>
struct Fold {
    static constexpr int M = N / 2;      // complex points per polynomial
    static constexpr int T = N / 8;      // threads per ciphertext
    static constexpr int LOG2M = __builtin_ctz(M);
    static constexpr bool ODD = (LOG2M & 1) != 0;   // one radix-2 pass needed
};

__device__ __forceinline__ void dft4(cplx x[4], bool inv)
{
    cplx t0 = cadd(x[0], x[2]);
    cplx t1 = csub(x[0], x[2]);
    cplx t2 = cadd(x[1], x[3]);
    cplx t3 = csub(x[1], x[3]);
    x[0] = cadd(t0, t2);
    x[2] = csub(t0, t2);
    cplx r = inv ? cmul_pi(t3) : cmul_mi(t3);
    x[1] = cadd(t1, r);
    x[3] = csub(t1, r);
}

// Twiddle then radix-4 butterfly for Stockham span Ns (inputs j + r*M/4).
template <int N, bool FULL = false>
__device__ __forceinline__ void r4_twiddle(cplx x[4], int k, int Ns, bool inv, const cplx *ztab)
{
    if (Ns > 1) {
        const int base = (Fold<N>::M / Ns) * k;     // zeta exponent of w^k, w = exp(-2 pi i/(4 Ns))
#pragma unroll
        for (int r = 1; r < 4; r++) {
            cplx w = zeta_pow<N, FULL>(ztab, inv ? -r * base : r * base);
            x[r] = cmul(x[r], w);
        }
    }
    dft4(x, inv);
}

// Forward transform of NP polynomials held in registers (v[p][r] = point
// t + r*T, t = thread index within the ciphertext's group); result left in
// buf[p*M + bin] in natural bin order.  All threads of the workgroup call it
// (it contains workgroup barriers).
template <int N, bool FULL = false, int NP = 2>
__device__ __forceinline__ void forward_pair(cplx (&v)[NP][4], cplx *buf, const cplx *ztab, int t)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    int Ns;
    if constexpr (Fold<N>::ODD) {
        // radix-2, span 1: butterflies j = t (points t, t+2T) and j = t+T (t+T, t+3T)
        cplx o[NP][4];
#pragma unroll
        for (int p = 0; p < NP; p++) {
            o[p][0] = cadd(v[p][0], v[p][2]);
            o[p][1] = csub(v[p][0], v[p][2]);
            o[p][2] = cadd(v[p][1], v[p][3]);
            o[p][3] = csub(v[p][1], v[p][3]);
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < NP; p++) {
            buf[p * M + 2 * t] = o[p][0];
            buf[p * M + 2 * t + 1] = o[p][1];
            buf[p * M + 2 * (t + T)] = o[p][2];
            buf[p * M + 2 * (t + T) + 1] = o[p][3];
        }
        __syncthreads();
        Ns = 2;
    } else {
#pragma unroll
        for (int p = 0; p < NP; p++)
            dft4(v[p], false);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                buf[p * M + 4 * t + r] = v[p][r];
        __syncthreads();
        Ns = 4;
    }
    for (; Ns < M; Ns *= 4) {
        cplx x[NP][4];
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                x[p][r] = buf[p * M + t + r * T];
        const int k = t & (Ns - 1);
#pragma unroll
        for (int p = 0; p < NP; p++)
            r4_twiddle<N, FULL>(x[p], k, Ns, false, ztab);
        __syncthreads();
        const int idxD = (t / Ns) * 4 * Ns + k;
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                buf[p * M + idxD + r * Ns] = x[p][r];
        __syncthreads();
    }
}

// Inverse (unnormalised) transform of buf[p*M + bin]; the result lands in
// registers, v[p][r] = point t + r*T.  Called by all threads of the workgroup.
template <int N, bool FULL = false, int NP = 2>
__device__ __forceinline__ void inverse_pair(cplx (&v)[NP][4], cplx *buf, const cplx *ztab, int t)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    constexpr int R4_END = Fold<N>::ODD ? M / 2 : M;   // radix-4 passes cover spans up to R4_END/4
    for (int Ns = 1; Ns < R4_END; Ns *= 4) {
        cplx x[NP][4];
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                x[p][r] = buf[p * M + t + r * T];
        const int k = t & (Ns - 1);
#pragma unroll
        for (int p = 0; p < NP; p++)
            r4_twiddle<N, FULL>(x[p], k, Ns, true, ztab);
        if (!Fold<N>::ODD && Ns * 4 == M) {
            // last pass: idxD == t, outputs at t + r*T
#pragma unroll
            for (int p = 0; p < NP; p++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    v[p][r] = x[p][r];
            return;
        }
        __syncthreads();
        const int idxD = (t / Ns) * 4 * Ns + k;
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                buf[p * M + idxD + r * Ns] = x[p][r];
        __syncthreads();
    }
    if constexpr (Fold<N>::ODD) {
        // radix-2, span M/2: butterflies j = t and j = t+T, inputs j and j+M/2
#pragma unroll
        for (int p = 0; p < NP; p++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int j = t + h * T;
                cplx x0 = buf[p * M + j];
                cplx x1 = buf[p * M + j + M / 2];
                x1 = cmul(x1, zeta_pow<N, FULL>(ztab, -4 * j));   // conj(w^j), w = exp(-2 pi i/M)
                v[p][h] = cadd(x0, x1);                    // point j
                v[p][h + 2] = csub(x0, x1);                // point j + M/2
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Wave-local transform for N in {512, 1024, 2048}: W = N/512 waves per
// ciphertext, each wave owning a 256-point region of the N/2-point transform.
// One radix-W stage crosses waves (a register butterfly plus one LDS
// exchange); the remaining four radix-4 passes run inside a wave with
// cross-lane register exchanges between them and no workgroup barrier.
// Bin b of the result is element b/W of region b mod W (see region_slot).
// ---------------------------------------------------------------------------
template <int N>
struct Wave {
    static constexpr int W = N / 512;
    static constexpr bool OK = (N >= 512 && N <= 2048);
};

// Position of element e of region k in LDS.  The XOR keeps every access
// pattern of the transform (contiguous reads, the digit-reversed output
// writes of the in-wave passes) and the MAC's gathers across regions free of
// bank conflicts.
template <int N>
__device__ __forceinline__ int region_slot(int k, int e)
{
    constexpr int W = Wave<N>::W;
    return k * 256 + (e ^ ((0xF5A0 >> (4 * ((e >> 4) & 3))) & 15) ^ (k * (16 / W)));
}

template <int N>
__device__ __forceinline__ int phys(int b)
{
    if constexpr (Wave<N>::OK)
        return region_slot<N>(b % Wave<N>::W, b / Wave<N>::W);
    else
        return b;
}

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// ---------------------------------------------------------------------------
// Cross-lane exchanges.  xswap<S>(a, b) exchanges lane bit log2(S) with the
// register pair (a, b): a lane with the bit clear ends with (its a, partner's
// a), a lane with it set with (partner's b, its b), the partner being lane
// l ^ S.  Four such swaps transpose a 4x4 block of (lane digit, register
// index), which moves the next radix-4 digit of the transform from the lanes
// into the registers without touching LDS.
//  S = 32, 16: v_permlane32_swap / v_permlane16_swap (gfx950), one
//              instruction per 32-bit word pair.
//  S = 8, 4:   two DPP moves with bank masks (the lanes of one half of each
//              row take the partner's word, the others keep their own).
//  S = 2, 1:   quad_perm DPP moves of both words and two selects.
// ---------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ void xswap_u32(uint32_t &a, uint32_t &b, bool hi)
{
    if constexpr (S == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (S == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (S == 8 || S == 4) {
        // DPP row_ror:n -- lane i of a row reads lane (i - n) mod 16.  Lanes
        // with the bit set (banks 2,3 for S = 8; banks 1,3 for S = 4) read
        // lane i - S; the others read lane i + S (row_ror:16-S).
        constexpr int FROM_LO = 0x120 + S, FROM_HI = 0x120 + 16 - S;
        constexpr int BM_HI = S == 8 ? 0xC : 0xA, BM_LO = S == 8 ? 0x3 : 0x5;
        const uint32_t na = (uint32_t)__builtin_amdgcn_update_dpp((int)a, (int)b, FROM_LO, 0xF, BM_HI, false);
        const uint32_t nb = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)a, FROM_HI, 0xF, BM_LO, false);
        a = na;
        b = nb;
    } else {
        static_assert(S == 2 || S == 1, "lane strides 1..32");
        // both partner words, then the selects: no select feeds a DPP move
        constexpr int QP = S == 1 ? 0xB1 : 0x4E;   // quad_perm [1,0,3,2] / [2,3,0,1]
        const uint32_t pa = (uint32_t)__builtin_amdgcn_mov_dpp((int)a, QP, 0xF, 0xF, true);
        const uint32_t pb = (uint32_t)__builtin_amdgcn_mov_dpp((int)b, QP, 0xF, 0xF, true);
        a = hi ? pb : a;
        b = hi ? b : pa;
    }
}

__device__ __forceinline__ uint32_t lo32(double d) { return (uint32_t)__double2loint(d); }
__device__ __forceinline__ uint32_t hi32(double d) { return (uint32_t)__double2hiint(d); }
__device__ __forceinline__ double mk64(uint32_t lo, uint32_t hi) { return __hiloint2double((int)hi, (int)lo); }

template <int S>
__device__ __forceinline__ void xswap(cplx &a, cplx &b, int l)
{
    const bool hi = (l & S) != 0;
    uint32_t a0 = lo32(a.re), a1 = hi32(a.re), a2 = lo32(a.im), a3 = hi32(a.im);
    uint32_t b0 = lo32(b.re), b1 = hi32(b.re), b2 = lo32(b.im), b3 = hi32(b.im);
    xswap_u32<S>(a0, b0, hi);
    xswap_u32<S>(a1, b1, hi);
    xswap_u32<S>(a2, b2, hi);
    xswap_u32<S>(a3, b3, hi);
    a = {mk64(a0, a1), mk64(a2, a3)};
    b = {mk64(b0, b1), mk64(b2, b3)};
}

// Transpose of the register index (2 bits) with the lane digit at lane
// strides (SH, SL): afterwards the lanes hold the former register digit.
template <int SH, int SL>
__device__ __forceinline__ void xtranspose(cplx (&x)[4], int l)
{
    xswap<SH>(x[0], x[2], l);
    xswap<SH>(x[1], x[3], l);
    xswap<SL>(x[0], x[1], l);
    xswap<SL>(x[2], x[3], l);
}

// Element of the 256-point region that holds transform output k0 + 4k1 +
// 16k2 + 64k3 after the in-wave passes: lane 16k0 + 4k1 + k2, register k3.
__device__ __forceinline__ int wave_bin(int l, int r)
{
    return (l >> 4) + 4 * ((l >> 2) & 3) + 16 * (l & 3) + 64 * r;
}

// In-wave 256-point transform of the regions reg[p]: four radix-4 passes,
// the data staying in registers between them.  Forward (decimation in
// frequency): lane l, register r start with point l + 64r; each pass is a
// register butterfly over the register digit followed by its twiddle, then a
// cross-lane transpose brings the next input digit into the registers; the
// outputs end at lane 16k0 + 4k1 + k2, register k3 (bin k0 + 4k1 + 16k2 +
// 64k3).  Inverse (decimation in time, unnormalised): the same steps in
// reverse order with conjugate twiddles.  If FROM_REGS the input is v[p][r]
// (forward) instead of LDS; if TO_REGS (inverse) the outputs (points l + 64r)
// stay in v instead of LDS.  LDS layout: element e of region rg of
// polynomial p at polys[p*M + region_slot(rg, e)].
template <int N, bool INV, bool FROM_REGS, bool TO_REGS, bool FULL = false, int NP = 2>
__device__ __forceinline__ void wave_passes(cplx (&v)[NP][4], cplx *polys, int rg, int l,
                                            const cplx *ztab)
{
    constexpr int M = N / 2;
    constexpr int E256 = M / 64;                 // zeta exponent of exp(-2 pi i/256)
    cplx x[NP][4];
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            x[p][r] = FROM_REGS ? v[p][r]
                                : polys[p * M + region_slot<N>(rg, INV ? wave_bin(l, r) : l + 64 * r)];
    // twiddle exponents of the three twiddled passes (units of exp(-2 pi i/256)):
    // pass 1 w256^(l k), pass 2 w64^((l & 15) k), pass 3 w16^((l & 3) k)
    const int tw[3] = {l, 4 * (l & 15), 16 * (l & 3)};
    auto twiddle = [&](int s) {
        cplx w[3];
#pragma unroll
        for (int k = 1; k < 4; k++)
            w[k - 1] = zeta_pow<N, FULL>(ztab, (INV ? -E256 : E256) * tw[s] * k);
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int k = 1; k < 4; k++)
                x[p][k] = cmul(x[p][k], w[k - 1]);
    };
    if constexpr (!INV) {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            dft4(x[p], false);
        }
        twiddle(0);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            xtranspose<32, 16>(x[p], l);
            dft4(x[p], false);
        }
        twiddle(1);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            xtranspose<8, 4>(x[p], l);
            dft4(x[p], false);
        }
        twiddle(2);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            xtranspose<2, 1>(x[p], l);
            dft4(x[p], false);
        }
        wave_sync();
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                polys[p * M + region_slot<N>(rg, wave_bin(l, r))] = x[p][r];
    } else {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            dft4(x[p], true);
            xtranspose<2, 1>(x[p], l);
        }
        twiddle(2);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            dft4(x[p], true);
            xtranspose<8, 4>(x[p], l);
        }
        twiddle(1);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            dft4(x[p], true);
            xtranspose<32, 16>(x[p], l);
        }
        twiddle(0);
#pragma unroll
        for (int p = 0; p < NP; p++)
            dft4(x[p], true);
        if constexpr (TO_REGS) {
#pragma unroll
            for (int p = 0; p < NP; p++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    v[p][r] = x[p][r];
        } else {
            wave_sync();
#pragma unroll
            for (int p = 0; p < NP; p++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    polys[p * M + region_slot<N>(rg, l + 64 * r)] = x[p][r];
        }
    }
}

// Forward transform of NP polynomials (v[p][r] = point t + r*T) into
// buf[p*M + phys(bin)].  Called by all threads of the workgroup.
template <int N, bool FULL = false, int NP = 2>
__device__ __forceinline__ void forward_wave(cplx (&v)[NP][4], cplx *buf, const cplx *ztab, int t)
{
    constexpr int M = N / 2, W = Wave<N>::W;
    const int wv = t >> 6, l = t & 63;
    if constexpr (W == 1) {
        __syncthreads();
        wave_passes<N, false, true, false, FULL, NP>(v, buf, 0, l, ztab);
    } else {
        // cross-wave DIF stage: z_k(n) = (sum_r x_{n + r M/W} w_W^{rk}) * w_M^{nk}
        cplx z[NP][4];
        int pos[4];
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if constexpr (W == 4) {
                cplx y[4] = {v[p][0], v[p][1], v[p][2], v[p][3]};
                dft4(y, false);                                 // y_k, n = t
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    z[p][k] = k ? cmul(y[k], zeta_pow<N, FULL>(ztab, 4 * t * k)) : y[k];
                    pos[k] = region_slot<N>(k, t);
                }
            } else {   // W == 2: points t, t+128 | t+256, t+384
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int n = t + 128 * h;
                    z[p][2 * h] = cadd(v[p][h], v[p][h + 2]);                          // k = 0
                    z[p][2 * h + 1] = cmul(csub(v[p][h], v[p][h + 2]), zeta_pow<N, FULL>(ztab, 4 * n));
                    pos[2 * h] = region_slot<N>(0, n);
                    pos[2 * h + 1] = region_slot<N>(1, n);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int k = 0; k < 4; k++)
                buf[p * M + pos[k]] = z[p][k];
        __syncthreads();
        wave_passes<N, false, false, false, FULL, NP>(v, buf, wv, l, ztab);
    }
    __syncthreads();
}

// Inverse (unnormalised) transform of buf[p*M + phys(bin)]; the result lands
// in registers, v[p][r] = point t + r*T.  Called by all threads.
template <int N, bool FULL = false, int NP = 2>
__device__ __forceinline__ void inverse_wave(cplx (&v)[NP][4], cplx *buf, const cplx *ztab, int t)
{
    constexpr int M = N / 2, W = Wave<N>::W;
    const int wv = t >> 6, l = t & 63;
    if constexpr (W == 1) {
        wave_passes<N, true, false, true, FULL, NP>(v, buf, 0, l, ztab);
    } else {
        wave_passes<N, true, false, false, FULL, NP>(v, buf, wv, l, ztab);
        __syncthreads();
        // cross-wave DIT stage: x_{n + r M/W} = sum_k (u_k(n) w_M^{-nk}) w_W^{-rk}
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if constexpr (W == 4) {
                cplx y[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const cplx u = buf[p * M + region_slot<N>(k, t)];
                    y[k] = k ? cmul(u, zeta_pow<N, FULL>(ztab, -4 * t * k)) : u;
                }
                dft4(y, true);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    v[p][r] = y[r];
            } else {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int n = t + 128 * h;
                    const cplx u0 = buf[p * M + region_slot<N>(0, n)];
                    const cplx u1 = cmul(buf[p * M + region_slot<N>(1, n)], zeta_pow<N, FULL>(ztab, -4 * n));
                    v[p][h] = cadd(u0, u1);
                    v[p][h + 2] = csub(u0, u1);
                }
            }
        }
    }
}

}  // namespace bslstm
