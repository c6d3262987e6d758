// gfx950 kernels of the encrypted spiking-LSTM timestep around the
// programmable bootstrap (nn_tfhe/Spike_LSTM.py:103-184 of the reference):
//
//   linear_extract_kernel   homomorphic Linear: plaintext weight polynomial x
//                           RLWE, coefficient-0 sample extraction, bias
//                           (Spike_LSTM.py:46-61), exact
//   trace_pack_kernel       LWE -> RLWE by modulus switch, trivial embedding
//                           and the log2(N)-step automorphism/key-switch
//                           trace (text_jax.py:268-345)
//   lut_product_kernel      plaintext LUT x RLWE products summed over the
//                           many-LUT slots (activation.py:103-106,
//                           Spike_LSTM.py:129-132)
//   pack_transpose_kernel,  packing of up to N LWEs into one RLWE with the
//   packing_mac_kernel,     packing key switching key (text_jax.py:211-239)
//   packing_finalize_kernel
//   poly_fft_kernel         negacyclic transform of integer polynomials into
//                           the reference's bin order (keys, LUTs)
//   torus_axpy_kernel,      sums of ciphertexts and X^e rotations
//   rotate_poly_kernel      (text_jax.py:27-43, 348-362)
//
// Torus values are int64 classes mod q = 2^log2_q, log2_q <= 64, stored as
// centred representatives; integer sums wrap mod 2^64, which is exact mod q.
// Products of a torus polynomial with a gadget digit or a LUT polynomial
// go through the folded FP64 transform of fold_fft.h, as the reference's
// jax_fourier products do; everything else is exact integer arithmetic.
#include <algorithm>

#include "capi_internal.h"
#include "fold_fft.h"
#include "torus_device.h"

namespace bslstm {

template <int N>
__device__ __forceinline__ void forward_any(cplx (&v)[2][4], cplx *buf, const cplx *ztab, int t)
{
    if constexpr (Wave<N>::OK)
        forward_wave<N>(v, buf, ztab, t);
    else
        forward_pair<N>(v, buf, ztab, t);
}

template <int N>
__device__ __forceinline__ void inverse_any(cplx (&v)[2][4], cplx *buf, const cplx *ztab, int t)
{
    if constexpr (Wave<N>::OK)
        inverse_wave<N>(v, buf, ztab, t);
    else
        inverse_pair<N>(v, buf, ztab, t);
}

template <int N>
__device__ __forceinline__ void forward_one(cplx (&v)[1][4], cplx *buf, const cplx *ztab, int t)
{
    if constexpr (Wave<N>::OK)
        forward_wave<N, false, 1>(v, buf, ztab, t);
    else
        forward_pair<N, false, 1>(v, buf, ztab, t);
}

// fold + twist of two real polynomials given by their coefficients
// lo[r] = p_{t+rT}, hi[r] = p_{t+rT+N/2}: u_k = (p_k - i p_{k+N/2}) zeta^k
template <int N>
__device__ __forceinline__ cplx fold_point(double lo, double hi, int t, int r, const cplx *ztab)
{
    return cmul({lo, -hi}, zeta_pow<N>(ztab, t + r * Fold<N>::T));
}

// inverse fold: coefficients k = t+rT (lo) and k+N/2 (hi) of an unnormalised
// inverse transform output w
template <int N>
__device__ __forceinline__ void unfold_point(cplx w, int t, int r, const cplx *ztab, double &lo,
                                             double &hi)
{
    const cplx z = cmul(w, zeta_pow<N>(ztab, -(t + r * Fold<N>::T)));
    constexpr double inv_m = 1.0 / (double)Fold<N>::M;
    lo = z.re * inv_m;
    hi = -z.im * inv_m;
}

// ---------------------------------------------------------------------------
// Homomorphic Linear (Spike_LSTM.py:27-61)
//
// For output row o with weight polynomial W_o (plaintext integers) and the
// input RLWE (A, B): mask_j = coefficient -j of W_o*A in Z[X]/(X^N+1), i.e.
//   mask_j = sum_i W_o[i] * E[(2N - j - i) mod 2N],  E = (A, -A) (the
//   negacyclic extension), body = sum_i W_o[i] * E_B[(2N - i) mod 2N] + bias_o.
// This is sample_extract(W_o*ct, 0) (text_jax.py:191-208) of the exact
// product; the reference forms the product with a rounded FP64 FFT
// (fft_product_twice, polynomial_jax.py:45-54).  Only the weight positions
// listed in `support` (the nonzero columns, uniform across the launch) are
// visited.  The weights of the workgroup's rows at a chunk of support
// positions are staged in LDS first (one broadcast read per position in the
// loop, instead of a dependent global load per position).
// ---------------------------------------------------------------------------
constexpr int LIN_THREADS = 256;
constexpr int LIN_OT = 8;       // output rows per workgroup
constexpr int LIN_CHUNK = 256;  // support positions staged per pass

__global__ __launch_bounds__(LIN_THREADS) void linear_extract_kernel(
    const int64_t *__restrict__ ct, int64_t ct_stride, const int64_t *__restrict__ w,
    const int32_t *__restrict__ support, int K, const int64_t *__restrict__ bias, int out_dim, int N,
    int lq, int64_t *__restrict__ masks, int64_t *__restrict__ bodies)
{
    extern __shared__ int64_t ext[];   // 2N: E[v] = A_v (v < N), -A_{v-N}
    __shared__ uint64_t wst[LIN_CHUNK][LIN_OT];   // W_o[i] of the chunk's positions (0: skipped)
    __shared__ int32_t sst[LIN_CHUNK];            // the chunk's positions i (0 where skipped)
    __shared__ uint64_t bst[LIN_CHUNK];           // body coefficient paired with W_o[i]
    const int b = blockIdx.z;
    const int o0 = blockIdx.y * LIN_OT;
    const bool body_block = blockIdx.x == 0;
    const int64_t *A = ct + (size_t)b * ct_stride;
    const int64_t *Bp = A + N;
    for (int k = threadIdx.x; k < N; k += LIN_THREADS) {
        const int64_t a = A[k];
        ext[k] = a;
        ext[N + k] = (int64_t)(0ULL - (uint64_t)a);
    }

    const int j = blockIdx.x * LIN_THREADS + threadIdx.x;
    const int ob = (int)threadIdx.x;   // body row of this thread (body block, ob < LIN_OT)
    uint64_t acc[LIN_OT];
#pragma unroll
    for (int oo = 0; oo < LIN_OT; oo++)
        acc[oo] = 0;
    uint64_t accb = (body_block && ob < LIN_OT && o0 + ob < out_dim && bias) ? (uint64_t)bias[o0 + ob] : 0ULL;
    for (int s0 = 0; s0 < K; s0 += LIN_CHUNK) {
        const int cn = min(LIN_CHUNK, K - s0);
        __syncthreads();   // ext written; the previous chunk consumed
        for (int ss = threadIdx.x; ss < cn; ss += LIN_THREADS) {
            const int i = support[s0 + ss];
            const bool ok = (unsigned)i < (unsigned)N;   // entries outside [0, N) are skipped
            sst[ss] = ok ? i : 0;
            // coefficient 0 of W*B: W_0 B_0 - sum_{i>0} W_i B_{N-i}
            if (body_block)
                bst[ss] = !ok ? 0ULL : (i == 0 ? (uint64_t)Bp[0] : 0ULL - (uint64_t)Bp[N - i]);
#pragma unroll
            for (int oo = 0; oo < LIN_OT; oo++)
                wst[ss][oo] = ok && o0 + oo < out_dim ? (uint64_t)w[(size_t)(o0 + oo) * N + i] : 0ULL;
        }
        __syncthreads();
        if (j < N) {
#pragma unroll 4
            for (int ss = 0; ss < cn; ss++) {
                const uint64_t av = (uint64_t)ext[(2 * N - j - sst[ss]) & (2 * N - 1)];
#pragma unroll
                for (int oo = 0; oo < LIN_OT; oo++)
                    acc[oo] += wst[ss][oo] * av;
            }
        }
        if (body_block && ob < LIN_OT) {
            for (int ss = 0; ss < cn; ss++)
                accb += wst[ss][ob] * bst[ss];
        }
    }
    if (j < N) {
#pragma unroll
        for (int oo = 0; oo < LIN_OT; oo++)
            if (o0 + oo < out_dim)
                masks[((size_t)b * out_dim + o0 + oo) * N + j] = tmod(acc[oo], lq);
    }
    if (body_block && ob < LIN_OT && o0 + ob < out_dim)
        bodies[(size_t)b * out_dim + o0 + ob] = tmod(accb, lq);
}

// ---------------------------------------------------------------------------
// Trace packing of one LWE into an RLWE (pack_lwe_to_rlwe, text_jax.py:314-345)
//
//   (a, b) <- round(x / N)                      modulus_switch to q/N
//   A = a_0 - sum_{i>0} a_{N-i} X^i, B = b      trivial_embedding (:268-287)
//   for step s < log2 N, k = 2^(log2 N - s) + 1:
//     (A', B') = (A(X^k), B(X^k))               apply_automorphism
//     (A_ks, B_ks) = sum_lev dec_lev(A') * G[s, lev]   (FFT-domain gadget
//                                                product with the Galois key)
//     A <- A + A_ks,  B <- B + B' + B_ks         (:290-311, :335-343)
// One workgroup per LWE; the RLWE stays in registers
// (coefficients t + i*N/8), the automorphism permutes it through LDS.
// gkey is (log2 N, l, 2, N/2) complex in the reference's bin order.
// ---------------------------------------------------------------------------
// slots (optional): LWE w uses the Galois key gkey + slots[w] * (log2 N * l *
// 2 * N/2), clamped to [0, slot_count) (batches that mix clients).
// RG = 2 (N in 512..2048): two groups of N/8 threads per LWE, each keeping one
// row of the RLWE and transforming every other pair of gadget levels, so a
// step runs on twice the threads; the launch uses it below 512 LWEs, where
// the LWEs alone leave CUs idle.  RG = 1: one group per LWE, two workgroups
// per CU -- from 512 LWEs at N = 512..2048 (4096 LWEs: 5.4 against 6.8 ms in
// the two-group form), and at every other N.  Both give bit-identical results.
template <int N, int RG>
__global__ __launch_bounds__(N / 8 * RG, (RG == 1 && Wave<N>::OK) ? 2 : 1) void trace_pack_kernel(
    const int64_t *__restrict__ lwe_a, const int64_t *__restrict__ lwe_b,
    const cplx *__restrict__ gkey, int l, int lb, int lq, int64_t *__restrict__ out,
    const int32_t *__restrict__ slots, int slot_count)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T, LOGN = __builtin_ctz(N);
    static_assert(RG == 1 || (RG == 2 && Wave<N>::OK), "two groups need the wave-local transform");
    constexpr int NR = 2 / RG;               // rows (A, B) a group keeps in registers
    // one group where a two-group form exists: same arithmetic as that form
    constexpr bool DUAL = RG == 1 && Wave<N>::OK;
    __shared__ cplx buf[RG * 2 * M];         // per group: two transform polynomials
    __shared__ cplx ztab[N];
    // 2N int64 for the permutation: aliases the last group's buffer, whose
    // first write after the reads below comes after forward_any's barrier
    int64_t *xb = reinterpret_cast<int64_t *>(buf + (RG - 1) * 2 * M);
    const int h = threadIdx.x / T;           // group
    const int t = threadIdx.x - h * T;
    cplx *mybuf = buf + h * 2 * M;
    const size_t wid = blockIdx.x;
    if (slots) {
        int sl = slots[wid];
        sl = sl < 0 ? 0 : (sl >= slot_count ? slot_count - 1 : sl);
        gkey += (size_t)sl * LOGN * l * 2 * M;
    }
    fill_root_table<N>(ztab);

    // X[c] = row (h*NR + c) of the RLWE: c = 0 is A, 1 is B for RG = 1
    uint64_t X[NR][8];
    {
        const int64_t *a = lwe_a + wid * N;
#pragma unroll
        for (int c = 0; c < NR; c++) {
            const bool isA = h * NR + c == 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int k = t + i * T;
                const int64_t ak = rhe_shift(tmod((uint64_t)a[k == 0 ? 0 : N - k], lq), LOGN);
                X[c][i] = isA ? (k == 0 ? (uint64_t)ak : 0ULL - (uint64_t)ak) : 0ULL;
            }
            if (!isA && t == 0)
                X[c][0] = (uint64_t)rhe_shift(tmod((uint64_t)lwe_b[wid], lq), LOGN);
        }
    }
    const double q = ldexp(1.0, lq), inv_q = ldexp(1.0, -lq);
    const int pairs = (l + 1) / 2;              // gadget levels are transformed two at a time
    const int iters = (pairs + RG - 1) / RG;    // forward transforms per group and step

    // L2 warm-up of each step's Galois key (l*2*N/2 complex, 640 KiB at
    // N = 2048, l = 20): at the start of the step the P workgroups resident on
    // one XCD (blockIdx.x mod 8) split its 128-byte lines and each thread
    // touches one word of up to two of them, so the multiply-accumulates that
    // follow the automorphism and the first transforms find the key in L2.
    // The words are folded into `warm` a step later and stored only under a
    // condition that never holds (lq <= 64).
    const int P = min(32, max(1, (int)((gridDim.x + 7) / 8)));
    const int lines = l * 2 * M / 8;
    const int lpw = (lines + P - 1) / P;
    const int line0 = ((int)(blockIdx.x >> 3) % P) * lpw;
    const int line_end = min(lines, line0 + lpw);
    uint32_t warm = 0, pf0 = 0, pf1 = 0;

#pragma unroll 1
    for (int st = 0; st < LOGN; st++) {
        const int kexp = (1 << (LOGN - st)) + 1;
        {
            warm ^= pf0 ^ pf1;
            pf0 = pf1 = 0;
            const uint32_t *cur = reinterpret_cast<const uint32_t *>(gkey + (size_t)st * l * 2 * M);
            const int j0 = line0 + (int)threadIdx.x, j1 = j0 + (int)blockDim.x;
            if (j0 < line_end)
                pf0 = cur[(size_t)j0 * 32];
            if (j1 < line_end)
                pf1 = cur[(size_t)j1 * 32];
        }
        // ---- automorphism X -> X^k (polynomial_jax.py:171-200)
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NR; c++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int j = t + i * T;
                const int e = (j * kexp) & (2 * N - 1);
                const int tg = e & (N - 1);
                xb[(h * NR + c) * N + tg] = (int64_t)(e >= N ? 0ULL - X[c][i] : X[c][i]);
            }
        __syncthreads();
        // r: digits source A(X^k); the B row takes B(X^k) now (B <- B + B(X^k)
        // + B_ks, the sum wraps mod 2^64)
        int64_t r[8];
        uint64_t cv[8];     // every level's carry bit of the class (carries_of)
#pragma unroll
        for (int i = 0; i < 8; i++) {
            r[i] = tmod((uint64_t)xb[t + i * T], lq);
            cv[i] = carries_of((uint64_t)r[i], lq);
        }
#pragma unroll
        for (int c = 0; c < NR; c++)
            if (h * NR + c == 1)
#pragma unroll
                for (int i = 0; i < 8; i++)
                    X[c][i] += (uint64_t)xb[N + t + i * T];
        // ---- gadget product with the Galois key of this step; group h
        // transforms the level pairs h, h + RG, ..., each pair's digits read
        // straight from bit fields of the class and its carry word
        // (digit_level / digit_at_level)
        const cplx *key = gkey + (size_t)st * l * 2 * M;
        // digits of level pair `pair`, their transforms, and the products with the
        // key rows of those levels added into A (in level order)
        auto pair_mac = [&](int pair, cplx (&A)[2][2][2]) {
            cplx v[2][4];
            const int lev = 2 * pair;
            double dg[2][8];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                // one uniform branch per level, none per coefficient
                const DigitLevel dl = digit_level(lev + c, l, lb, lq);
                if (dl.kind == 1) {
#pragma unroll
                    for (int i = 0; i < 8; i++)
                        dg[c][i] = (double)digit_at_level<1>(dl, (uint64_t)r[i], cv[i]);
                } else if (dl.kind == 2) {
#pragma unroll
                    for (int i = 0; i < 8; i++)
                        dg[c][i] = (double)digit_at_level<2>(dl, (uint64_t)r[i], cv[i]);
                } else {
#pragma unroll
                    for (int i = 0; i < 8; i++)
                        dg[c][i] = 0.0;
                }
            }
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++)
                    v[c][rr] = fold_point<N>(dg[c][rr], dg[c][rr + 4], t, rr, ztab);
            forward_any<N>(v, mybuf, ztab, t);
#pragma unroll
            for (int pp = 0; pp < 2; pp++) {
                const int rb = t + pp * T;   // bin pair (rb, M-1-rb) = F-bins 2rb, 2rb+1
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (lev + c >= l)
                        break;
                    const cplx d0 = mybuf[c * M + phys<N>(rb)];
                    const cplx d1 = mybuf[c * M + phys<N>(M - 1 - rb)];
#pragma unroll
                    for (int a = 0; a < 2; a++) {
                        const cplx *kb = key + ((size_t)(lev + c) * 2 + a) * M;
                        A[a][pp][0] = cfma(d0, kb[2 * rb], A[a][pp][0]);
                        A[a][pp][1] = cfma_conj(kb[2 * rb + 1], d1, A[a][pp][1]);
                    }
                }
            }
        };
        // acc sums the pairs h, h + 2, ... (two groups) or all pairs (one
        // group); with DUAL one group keeps the even and odd pairs' sums apart
        // (acc, acc1) and adds them at the end, which is the two-group form's
        // arithmetic exactly
        cplx acc[2][2][2], acc1[2][2][2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int pp = 0; pp < 2; pp++) {
                acc[a][pp][0] = acc[a][pp][1] = {0.0, 0.0};
                acc1[a][pp][0] = acc1[a][pp][1] = {0.0, 0.0};
            }
        if constexpr (DUAL) {
#pragma unroll 1
            for (int it = 0; it < (pairs + 1) / 2; it++) {
                pair_mac(2 * it, acc);
                if (2 * it + 1 < pairs)
                    pair_mac(2 * it + 1, acc1);
            }
        } else {
#pragma unroll 1
            for (int it = 0; it < iters; it++)
                pair_mac(it * RG + h, acc);
        }
        __syncthreads();
        if constexpr (RG == 1) {
#pragma unroll
            for (int pp = 0; pp < 2; pp++) {
                const int rb = t + pp * T;
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    buf[a * M + phys<N>(rb)] = DUAL ? cadd(acc[a][pp][0], acc1[a][pp][0]) : acc[a][pp][0];
                    buf[a * M + phys<N>(M - 1 - rb)] =
                        DUAL ? cadd(acc[a][pp][1], acc1[a][pp][1]) : acc[a][pp][1];
                }
            }
        } else {
            // group h hands its partial product for row 1-h to the other group
            // (second polynomial of that group's buffer), then sums row h
            // (component selects, not acc[h]: a per-thread index, or a select
            // of whole elements, would put acc in memory)
            const bool g1 = h != 0;
            auto pick = [](bool sel, const cplx &x1, const cplx &x0) -> cplx {
                return {sel ? x1.re : x0.re, sel ? x1.im : x0.im};
            };
#pragma unroll
            for (int pp = 0; pp < 2; pp++) {
                const int rb = t + pp * T;
                cplx *ob = buf + (1 - h) * 2 * M + M;
                ob[phys<N>(rb)] = pick(g1, acc[0][pp][0], acc[1][pp][0]);
                ob[phys<N>(M - 1 - rb)] = pick(g1, acc[0][pp][1], acc[1][pp][1]);
            }
            __syncthreads();
#pragma unroll
            for (int pp = 0; pp < 2; pp++) {
                const int rb = t + pp * T;
                const cplx own0 = pick(g1, acc[1][pp][0], acc[0][pp][0]);
                const cplx own1 = pick(g1, acc[1][pp][1], acc[0][pp][1]);
                mybuf[phys<N>(rb)] = cadd(own0, mybuf[M + phys<N>(rb)]);
                mybuf[phys<N>(M - 1 - rb)] = cadd(own1, mybuf[M + phys<N>(M - 1 - rb)]);
            }
        }
        __syncthreads();
        cplx wv[NR][4];
        if constexpr (Wave<N>::OK)
            inverse_wave<N, false, NR>(wv, mybuf, ztab, t);
        else
            inverse_pair<N, false, NR>(wv, mybuf, ztab, t);
#pragma unroll
        for (int c = 0; c < NR; c++) {
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                double lo, hi;
                unfold_point<N>(wv[c][rr], t, rr, ztab, lo, hi);
                X[c][rr] += (uint64_t)f64_to_torus(lo, q, inv_q, lq);
                X[c][rr + 4] += (uint64_t)f64_to_torus(hi, q, inv_q, lq);
            }
        }
    }
    int64_t *dst = out + wid * 2 * N + (size_t)h * NR * N;
    if ((warm ^ pf0 ^ pf1) == 0x9E3779B9u && lq > 64)
        dst[t] = (int64_t)warm;
#pragma unroll
    for (int c = 0; c < NR; c++)
#pragma unroll
        for (int i = 0; i < 8; i++)
            dst[c * N + t + i * T] = tmod(X[c][i], lq);
}

// ---------------------------------------------------------------------------
// Plaintext-LUT products (prepare_LUT, activation.py:103-106, summed over the
// many-LUT slots as Spike_LSTM.py:129-132 does), for a batch of sequences
// sharing the LUTs:
//   out[s, g] = sum_k round(LUT[k, g](X) * RLWE[s, k, g])   (both halves), mod q
// lut_f holds the LUT polynomials already transformed (reference bin order).
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(N / 8) void lut_product_kernel(const int64_t *__restrict__ rlwe,
                                                            const cplx *__restrict__ lut_f,
                                                            int64_t G, int n_lut, int lq,
                                                            int64_t *__restrict__ out)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    __shared__ cplx buf[2 * M];
    __shared__ cplx ztab[N];
    const int t = threadIdx.x;
    const int64_t g = blockIdx.x;
    const int64_t sq = blockIdx.y;
    fill_root_table<N>(ztab);
    const double q = ldexp(1.0, lq), inv_q = ldexp(1.0, -lq);
    uint64_t S[2][8];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int i = 0; i < 8; i++)
            S[a][i] = 0;
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < n_lut; k++) {
        const int64_t idx = (int64_t)k * G + g;
        const int64_t *src = rlwe + ((sq * n_lut + k) * G + g) * 2 * N;
        cplx v[2][4];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
                v[a][rr] = fold_point<N>((double)src[a * N + t + rr * T],
                                         (double)src[a * N + t + (rr + 4) * T], t, rr, ztab);
        forward_any<N>(v, buf, ztab, t);
        const cplx *L = lut_f + idx * M;
#pragma unroll
        for (int pp = 0; pp < 2; pp++) {
            const int rb = t + pp * T;
            const cplx l0 = L[2 * rb], l1 = cconj(L[2 * rb + 1]);
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int p0 = a * M + phys<N>(rb), p1 = a * M + phys<N>(M - 1 - rb);
                buf[p0] = cmul(buf[p0], l0);
                buf[p1] = cmul(buf[p1], l1);
            }
        }
        __syncthreads();
        cplx wv[2][4];
        inverse_any<N>(wv, buf, ztab, t);
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                double lo, hi;
                unfold_point<N>(wv[a][rr], t, rr, ztab, lo, hi);
                S[a][rr] += (uint64_t)f64_to_torus(lo, q, inv_q, lq);
                S[a][rr + 4] += (uint64_t)f64_to_torus(hi, q, inv_q, lq);
            }
    }
    int64_t *dst = out + (sq * G + g) * 2 * N;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int i = 0; i < 8; i++)
            dst[a * N + t + i * T] = tmod(S[a][i], lq);
}

// ---------------------------------------------------------------------------
// Negacyclic transform of integer polynomials, output in the reference's
// bin order F_j = p(zeta^(2j+1)), j < N/2 (jax_fourier, polynomial_jax.py:11-15).
// Two polynomials per workgroup.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(N / 8) void poly_fft_kernel(const int64_t *__restrict__ polys,
                                                         int64_t cnt, cplx *__restrict__ out)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    __shared__ cplx buf[2 * M];
    __shared__ cplx ztab[N];
    const int t = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * 2;
    fill_root_table<N>(ztab);
    __syncthreads();
    cplx v[2][4];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const bool ok = p0 + c < cnt;
        const int64_t *src = polys + (p0 + c) * N;
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
            v[c][rr] = ok ? fold_point<N>((double)src[t + rr * T], (double)src[t + (rr + 4) * T],
                                          t, rr, ztab)
                          : cplx{0.0, 0.0};
    }
    forward_any<N>(v, buf, ztab, t);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        if (p0 + c >= cnt)
            break;
        cplx *dst = out + (p0 + c) * M;
#pragma unroll
        for (int pp = 0; pp < 2; pp++) {
            const int rb = t + pp * T;
            dst[2 * rb] = buf[c * M + phys<N>(rb)];
            dst[2 * rb + 1] = cconj(buf[c * M + phys<N>(M - 1 - rb)]);
        }
    }
}

// Inverse of poly_fft_kernel: spectra in the reference bin order
// (F_j = p(zeta^(2j+1)), j < N/2) back to the coefficients, rounded to the
// nearest integer (jax_inversefourier's rounding, polynomial_jax.py:17-21).
// Two polynomials per workgroup, as the forward kernel.
template <int N>
__global__ __launch_bounds__(N / 8) void poly_ifft_kernel(const cplx *__restrict__ spectra,
                                                          int64_t cnt, int64_t *__restrict__ out)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    __shared__ cplx buf[2 * M];
    __shared__ cplx ztab[N];
    const int t = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * 2;
    fill_root_table<N>(ztab);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const bool ok = p0 + c < cnt;
        const cplx *src = spectra + (p0 + c) * M;
#pragma unroll
        for (int pp = 0; pp < 2; pp++) {
            const int rb = t + pp * T;
            buf[c * M + phys<N>(rb)] = ok ? src[2 * rb] : cplx{0.0, 0.0};
            buf[c * M + phys<N>(M - 1 - rb)] = ok ? cconj(src[2 * rb + 1]) : cplx{0.0, 0.0};
        }
    }
    __syncthreads();
    cplx w[2][4];
    inverse_any<N>(w, buf, ztab, t);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        if (p0 + c >= cnt)
            break;
        int64_t *dst = out + (p0 + c) * N;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            double lo, hi;
            unfold_point<N>(w[c][rr], t, rr, ztab, lo, hi);
            dst[t + rr * T] = (int64_t)rint(lo);
            dst[t + (rr + 4) * T] = (int64_t)rint(hi);
        }
    }
}

// ---------------------------------------------------------------------------
// Packing of n_lwes LWEs (masks of length n_in) into one RLWE
// (packing, text_jax.py:211-239):
//   row j (j < n_in) is the polynomial P_j with coefficient i = a_i[j]
//   (i < n_lwes, zero above; only the n_lwes nonzero coefficients are
//   stored);
//   A = -sum_j sum_lev dec_lev(P_j) * KSP[j, lev, 0]
//   B = (b_0 .. b_{n_lwes-1}, 0 ..) - sum_j sum_lev dec_lev(P_j) * KSP[j, lev, 1]
// 1. pack_transpose_kernel: masks -> rows P_j (coalesced tile transpose)
// 2. packing_mac_kernel: each workgroup transforms the digits of a block of
//    rows and accumulates digit x key per bin; one partial sum per block
// 3. packing_reduce_kernel sums the partials across the chip in a fixed order;
//    packing_finalize_kernel: one inverse transform per half, rounding, body.
// The sum over rows is rounded once, after the transform domain sum (the
// reference rounds per row, then sums).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pack_transpose_kernel(const int64_t *__restrict__ masks,
                                                             int n_lwes, int n_in, int N, int lq,
                                                             int64_t *__restrict__ rows)
{
    __shared__ int64_t tile[64][65];
    const int p = blockIdx.z;
    const int i0 = blockIdx.y * 64;   // coefficient (LWE index)
    const int j0 = blockIdx.x * 64;   // row (mask coordinate)
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int ii = e >> 6, jj = e & 63;
        const int i = i0 + ii, j = j0 + jj;
        tile[ii][jj] = (i < n_lwes && j < n_in)
                           ? tmod((uint64_t)masks[((size_t)p * n_lwes + i) * n_in + j], lq)
                           : 0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int jj = e >> 6, ii = e & 63;
        const int i = i0 + ii, j = j0 + jj;
        if (j < n_in && i < n_lwes)
            rows[((size_t)p * n_in + j) * n_lwes + i] = tile[ii][jj];
    }
}

// slots (optional): packing p uses the key ksk + slots[p] * (n_in * l * 2 *
// N/2), clamped to [0, slot_count) (batches that mix clients).
template <int N>
__global__ __launch_bounds__(N / 8) void packing_mac_kernel(const int64_t *__restrict__ rows,
                                                            int n_lwes, int n_in, int rows_per_block,
                                                            const cplx *__restrict__ ksk, int l,
                                                            int lb, int lq,
                                                            cplx *__restrict__ partial,
                                                            const int32_t *__restrict__ slots,
                                                            int slot_count)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    __shared__ cplx buf[M];   // one transform at a time: 48 KiB of LDS, three workgroups per CU
    __shared__ cplx ztab[N];
    const int t = threadIdx.x;
    const int p = blockIdx.y, blk = blockIdx.x, NB = gridDim.x;
    if (slots) {
        int sl = slots[p];
        sl = sl < 0 ? 0 : (sl >= slot_count ? slot_count - 1 : sl);
        ksk += (size_t)sl * n_in * l * 2 * M;
    }
    fill_root_table<N>(ztab);
    __syncthreads();
    const int j0 = blk * rows_per_block;
    const int j1 = min(n_in, j0 + rows_per_block);
    cplx acc[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int pp = 0; pp < 2; pp++)
            acc[a][pp][0] = acc[a][pp][1] = {0.0, 0.0};
#pragma unroll 1
    for (int j = j0; j < j1; j++) {
        const int64_t *row = rows + ((size_t)p * n_in + j) * n_lwes;
        int64_t r[8];
#pragma unroll
        for (int i = 0; i < 8; i++)
            r[i] = t + i * T < n_lwes ? row[t + i * T] : 0;
#pragma unroll 1
        for (int lev = 0; lev < l; lev++) {
            cplx v[1][4];
            double dg[8];
#pragma unroll
            for (int i = 0; i < 8; i++)
                dg[i] = (double)next_digit(r[i], lq - (lev + 1) * lb);
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
                v[0][rr] = fold_point<N>(dg[rr], dg[rr + 4], t, rr, ztab);
            forward_one<N>(v, buf, ztab, t);
#pragma unroll
            for (int pp = 0; pp < 2; pp++) {
                const int rb = t + pp * T;
                const cplx d0 = buf[phys<N>(rb)];
                const cplx d1 = buf[phys<N>(M - 1 - rb)];
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    const cplx *kb = ksk + (((size_t)j * l + lev) * 2 + a) * M;
                    acc[a][pp][0] = cfma(d0, kb[2 * rb], acc[a][pp][0]);
                    acc[a][pp][1] = cfma_conj(kb[2 * rb + 1], d1, acc[a][pp][1]);
                }
            }
        }
    }
    // partial (P, NB, 2, M), U-domain bins in natural order
    cplx *dst = partial + ((size_t)p * NB + blk) * 2 * M;
#pragma unroll
    for (int pp = 0; pp < 2; pp++) {
        const int rb = t + pp * T;
#pragma unroll
        for (int a = 0; a < 2; a++) {
            dst[a * M + rb] = acc[a][pp][0];
            dst[a * M + M - 1 - rb] = acc[a][pp][1];
        }
    }
}

// Sum of the per-block partial spectra across the chip, in an order fixed by
// NB alone (so a packing gives the same sum in any batch): the blocks are cut
// into RED_G runs; one thread per (bin, run) adds its run in block order, and
// the run sums are then added in run order.
constexpr int RED_G = 8;    // runs of blocks per bin
constexpr int RED_B = 32;   // bins per workgroup (RED_G * RED_B = 256 threads)

__global__ __launch_bounds__(RED_G * RED_B) void packing_reduce_kernel(const cplx *__restrict__ partial,
                                                                      int NB, int M,
                                                                      cplx *__restrict__ reduced)
{
    __shared__ cplx run_sum[RED_G][RED_B];
    const int bl = (int)threadIdx.x % RED_B, g = (int)threadIdx.x / RED_B;
    const int bin = blockIdx.x * RED_B + bl;
    const int a = blockIdx.y;
    const size_t p = blockIdx.z;
    const int per = (NB + RED_G - 1) / RED_G;
    const int b0 = g * per, b1 = min(NB, b0 + per);
    cplx sum = {0.0, 0.0};
    if (bin < M)
        for (int blk = b0; blk < b1; blk++)
            sum = cadd(sum, partial[((p * NB + blk) * 2 + a) * M + bin]);
    run_sum[g][bl] = sum;
    __syncthreads();
    if (g == 0 && bin < M) {
        cplx tot = run_sum[0][bl];
#pragma unroll
        for (int r = 1; r < RED_G; r++)
            tot = cadd(tot, run_sum[r][bl]);
        reduced[(p * 2 + a) * M + bin] = tot;
    }
}

template <int N>
__global__ __launch_bounds__(N / 8) void packing_finalize_kernel(const cplx *__restrict__ partial,
                                                                 int NB,
                                                                 const int64_t *__restrict__ bodies,
                                                                 int n_lwes, int lq,
                                                                 int64_t *__restrict__ out)
{
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    __shared__ cplx buf[2 * M];
    __shared__ cplx ztab[N];
    const int t = threadIdx.x;
    const int p = blockIdx.x;
    fill_root_table<N>(ztab);
    const cplx *src = partial + (size_t)p * NB * 2 * M;
#pragma unroll
    for (int pp = 0; pp < 2; pp++) {
        const int rb = t + pp * T;
#pragma unroll
        for (int a = 0; a < 2; a++) {
            cplx s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
            for (int blk = 0; blk < NB; blk++) {
                s0 = cadd(s0, src[((size_t)blk * 2 + a) * M + rb]);
                s1 = cadd(s1, src[((size_t)blk * 2 + a) * M + M - 1 - rb]);
            }
            buf[a * M + phys<N>(rb)] = s0;
            buf[a * M + phys<N>(M - 1 - rb)] = s1;
        }
    }
    __syncthreads();
    cplx wv[2][4];
    inverse_any<N>(wv, buf, ztab, t);
    const double q = ldexp(1.0, lq), inv_q = ldexp(1.0, -lq);
    int64_t *dst = out + (size_t)p * 2 * N;
    const int64_t *bp = bodies + (size_t)p * n_lwes;
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        double lo, hi;
        unfold_point<N>(wv[0][rr], t, rr, ztab, lo, hi);
        const int k0 = t + rr * Fold<N>::T, k1 = k0 + N / 2;
        dst[k0] = tmod(0ULL - (uint64_t)f64_to_torus(lo, q, inv_q, lq), lq);
        dst[k1] = tmod(0ULL - (uint64_t)f64_to_torus(hi, q, inv_q, lq), lq);
        unfold_point<N>(wv[1][rr], t, rr, ztab, lo, hi);
        const uint64_t b0 = k0 < n_lwes ? (uint64_t)bp[k0] : 0ULL;
        const uint64_t b1 = k1 < n_lwes ? (uint64_t)bp[k1] : 0ULL;
        dst[N + k0] = tmod(b0 - (uint64_t)f64_to_torus(lo, q, inv_q, lq), lq);
        dst[N + k1] = tmod(b1 - (uint64_t)f64_to_torus(hi, q, inv_q, lq), lq);
    }
}

// ---------------------------------------------------------------------------
// Ciphertext arithmetic: out = centred(sx*x + sy*y) (sum_ciphertext_ciphertext,
// text_jax.py:27-43) and out = X^e * in per polynomial (coef_rotation,
// polynomial_jax.py:150-169; rotate_ciphertext, text_jax.py:348-362).
// ---------------------------------------------------------------------------
__global__ void torus_axpy_kernel(const int64_t *__restrict__ x, const int64_t *__restrict__ y,
                                  int64_t n, int64_t sx, int64_t sy, int lq,
                                  int64_t *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = tmod((uint64_t)sx * (uint64_t)x[i] + (uint64_t)sy * (uint64_t)y[i], lq);
}

__global__ void rotate_poly_kernel(const int64_t *__restrict__ in, int64_t rows, int N,
                                   const int64_t *__restrict__ exps, int64_t exp_every, int lq,
                                   int64_t *__restrict__ out)
{
    const int64_t row = blockIdx.y;
    const int64_t e = exps[row / exp_every];
    const int twoN = 2 * N;
    const int al = (int)(((e % twoN) + twoN) % twoN);
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < N; s += gridDim.x * blockDim.x) {
        // (X^al * p)[s] = p[w] if w < N else -p[w - N], w = (s - al) mod 2N
        int w = s - al;
        if (w < 0)
            w += twoN;
        const uint64_t v = (uint64_t)(w < N ? in[row * N + w] : in[row * N + w - N]);
        out[row * N + s] = tmod(w < N ? v : 0ULL - v, lq);
    }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
#define BSLSTM_DEGREE_SWITCH(deg, CALL)                                                          \
    switch (deg) {                                                                               \
    case 128: { constexpr int NN = 128; CALL; } break;                                           \
    case 256: { constexpr int NN = 256; CALL; } break;                                           \
    case 512: { constexpr int NN = 512; CALL; } break;                                           \
    case 1024: { constexpr int NN = 1024; CALL; } break;                                         \
    case 2048: { constexpr int NN = 2048; CALL; } break;                                         \
    case 4096: { constexpr int NN = 4096; CALL; } break;                                         \
    default: return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");                         \
    }

static bool degree_ok(int64_t d) { return d >= 128 && d <= 4096 && is_pow2(d); }

static int check_gadget(int64_t l, int64_t log2_beta, int64_t log2_q)
{
    if (log2_q < 2 || log2_q > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "log2_q must be in [2, 64]");
    if (l < 1 || log2_beta < 1 || log2_beta > 30 || l * log2_beta > log2_q)
        return fail(BSLSTM_INVALID_ARGUMENT, "need l >= 1, 1 <= log2_beta <= 30, l*log2_beta <= log2_q");
    return BSLSTM_OK;
}

}  // namespace bslstm

using namespace bslstm;

extern "C" int bslstm_linear_extract(void *stream, const int64_t *ct, int64_t ct_elems,
                                     const int64_t *weights, const int32_t *support, int64_t K,
                                     const int64_t *bias, int64_t batch, int64_t out_dim,
                                     int64_t degree, int64_t log2_q, int64_t *masks, int64_t *bodies)
{
    clear_error();
    if (!is_pow2(degree) || degree < 2 || degree > 4096)
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (log2_q < 2 || log2_q > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "log2_q must be in [2, 64]");
    if (batch < 0 || batch > 65535 || out_dim < 0 || K < 0 || K > degree)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad batch/out_dim/K");
    int64_t stride;
    if (ct_elems == 2 * degree)
        stride = 0;
    else if (ct_elems == batch * 2 * degree)
        stride = 2 * degree;
    else
        return fail(BSLSTM_INVALID_ARGUMENT, "ct has %lld elements; expected 2N or batch*2N",
                    (long long)ct_elems);
    if (batch == 0 || out_dim == 0)
        return BSLSTM_OK;
    if (!ct || !weights || (K > 0 && !support) || !masks || !bodies)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const int N = (int)degree;
    dim3 grid((unsigned)((N + LIN_THREADS - 1) / LIN_THREADS),
              (unsigned)((out_dim + LIN_OT - 1) / LIN_OT), (unsigned)batch);
    if (grid.y > 65535)
        return fail(BSLSTM_INVALID_ARGUMENT, "out_dim too large");
    hipLaunchKernelGGL(linear_extract_kernel, grid, dim3(LIN_THREADS), sizeof(int64_t) * (2 * N),
                       (hipStream_t)stream, ct, stride, weights, support, (int)K, bias,
                       (int)out_dim, N, (int)log2_q, masks, bodies);
    return launch_status("linear extract");
}

static int pack_lwe_run(void *stream, const int64_t *lwe_a, const int64_t *lwe_b, int64_t count,
                        const double *galois_key_fft, const int32_t *slots, int64_t slot_count,
                        int64_t degree, int64_t log2_beta, int64_t l, int64_t log2_q, int64_t *out)
{
    clear_error();
    if (!degree_ok(degree))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (int s = check_gadget(l, log2_beta, log2_q))
        return s;
    if (count < 0 || count > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad count");
    if (count == 0)
        return BSLSTM_OK;
    if (!lwe_a || !lwe_b || !galois_key_fft || !out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const cplx *gk = reinterpret_cast<const cplx *>(galois_key_fft);
    // two thread groups per LWE for latency (fewer LWEs than two per CU);
    // one group, two workgroups per CU, for throughput -- both forms give the
    // same result for every LWE, so it does not depend on the batch
    const bool two = degree >= 512 && degree <= 2048 && count < 512;
    if (two) {
        BSLSTM_DEGREE_SWITCH(degree, hipLaunchKernelGGL((trace_pack_kernel<NN, Wave<NN>::OK ? 2 : 1>),
                                                        dim3((unsigned)count),
                                                        dim3(NN / 8 * (Wave<NN>::OK ? 2 : 1)), 0,
                                                        (hipStream_t)stream, lwe_a, lwe_b, gk,
                                                        (int)l, (int)log2_beta, (int)log2_q, out,
                                                        slots, (int)slot_count));
    } else {
        BSLSTM_DEGREE_SWITCH(degree, hipLaunchKernelGGL((trace_pack_kernel<NN, 1>), dim3((unsigned)count),
                                                        dim3(NN / 8), 0, (hipStream_t)stream, lwe_a,
                                                        lwe_b, gk, (int)l, (int)log2_beta,
                                                        (int)log2_q, out, slots, (int)slot_count));
    }
    return launch_status("trace packing");
}

extern "C" int bslstm_pack_lwe_to_rlwe(void *stream, const int64_t *lwe_a, const int64_t *lwe_b,
                                       int64_t count, const double *galois_key_fft,
                                       int64_t degree, int64_t log2_beta, int64_t l,
                                       int64_t log2_q, int64_t *out)
{
    return pack_lwe_run(stream, lwe_a, lwe_b, count, galois_key_fft, nullptr, 1, degree, log2_beta,
                        l, log2_q, out);
}

extern "C" int bslstm_pack_lwe_to_rlwe_slots(void *stream, const int64_t *lwe_a,
                                             const int64_t *lwe_b, int64_t count,
                                             const double *galois_key_pool, const int32_t *slots,
                                             int64_t slot_count, int64_t degree, int64_t log2_beta,
                                             int64_t l, int64_t log2_q, int64_t *out)
{
    if (!slots || slot_count < 1 || slot_count > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "slots and 1 <= slot_count required");
    return pack_lwe_run(stream, lwe_a, lwe_b, count, galois_key_pool, slots, slot_count, degree,
                        log2_beta, l, log2_q, out);
}

extern "C" int bslstm_lut_product_sum(void *stream, const int64_t *rlwe, const double *lut_fft,
                                      int64_t batch, int64_t groups, int64_t n_lut, int64_t degree,
                                      int64_t log2_q, int64_t *out)
{
    clear_error();
    if (!degree_ok(degree))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (log2_q < 2 || log2_q > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "log2_q must be in [2, 64]");
    if (groups < 0 || groups > 0x7fffffffLL || n_lut < 1 || batch < 0 || batch > 65535)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad batch/groups/n_lut (batch <= 65535)");
    if (groups == 0 || batch == 0)
        return BSLSTM_OK;
    if (!rlwe || !lut_fft || !out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const cplx *lf = reinterpret_cast<const cplx *>(lut_fft);
    BSLSTM_DEGREE_SWITCH(degree, hipLaunchKernelGGL((lut_product_kernel<NN>),
                                                    dim3((unsigned)groups, (unsigned)batch),
                                                    dim3(NN / 8), 0, (hipStream_t)stream, rlwe, lf,
                                                    groups, (int)n_lut, (int)log2_q, out));
    return launch_status("LUT product");
}

extern "C" int bslstm_poly_fft(void *stream, const int64_t *polys, int64_t count, int64_t degree,
                               double *out)
{
    clear_error();
    if (!degree_ok(degree))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (count < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad count");
    if (count == 0)
        return BSLSTM_OK;
    if (!polys || !out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const int64_t blocks = (count + 1) / 2;
    if (blocks > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "count too large");
    cplx *o = reinterpret_cast<cplx *>(out);
    BSLSTM_DEGREE_SWITCH(degree, hipLaunchKernelGGL((poly_fft_kernel<NN>), dim3((unsigned)blocks),
                                                    dim3(NN / 8), 0, (hipStream_t)stream, polys,
                                                    count, o));
    return launch_status("polynomial transform");
}

extern "C" int bslstm_poly_ifft(void *stream, const double *spectra, int64_t count, int64_t degree,
                                int64_t *out)
{
    clear_error();
    if (!degree_ok(degree))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (count < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad count");
    if (count == 0)
        return BSLSTM_OK;
    if (!spectra || !out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const int64_t blocks = (count + 1) / 2;
    if (blocks > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "count too large");
    const cplx *sp = reinterpret_cast<const cplx *>(spectra);
    BSLSTM_DEGREE_SWITCH(degree, hipLaunchKernelGGL((poly_ifft_kernel<NN>), dim3((unsigned)blocks),
                                                    dim3(NN / 8), 0, (hipStream_t)stream, sp,
                                                    count, out));
    return launch_status("inverse polynomial transform");
}

// Row blocks of the packing MAC: fixed by n_in alone, never by the number of
// packings, so that a packing's sum order (and result) does not depend on the
// batch it runs in.
static int packing_blocks(int64_t n_in) { return (int)std::min<int64_t>(n_in, 256); }

extern "C" int64_t bslstm_packing_workspace_bytes(int64_t packs, int64_t n_in, int64_t degree)
{
    if (packs < 0 || n_in < 1 || !is_pow2(degree))
        return -1;
    const int64_t nb = packing_blocks(n_in);
    return packs * n_in * degree * 8 + packs * nb * 2 * (degree / 2) * 16;
}

static int packing_run(void *stream, const int64_t *masks, const int64_t *bodies, int64_t packs,
                       int64_t n_lwes, int64_t n_in, const double *ksk_fft, const int32_t *slots,
                       int64_t slot_count, int64_t degree, int64_t log2_beta, int64_t l,
                       int64_t log2_q, void *workspace, int64_t workspace_bytes, int64_t *out)
{
    clear_error();
    if (!degree_ok(degree))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (int s = check_gadget(l, log2_beta, log2_q))
        return s;
    if (packs < 0 || packs > 65535 || n_lwes < 0 || n_lwes > degree || n_in < 1 || n_in > (1 << 24))
        return fail(BSLSTM_INVALID_ARGUMENT, "bad packs/n_lwes/n_in (n_lwes <= degree)");
    if (packs == 0)
        return BSLSTM_OK;
    const int64_t need = bslstm_packing_workspace_bytes(packs, n_in, degree);
    if (workspace_bytes < need)
        return fail(BSLSTM_INVALID_ARGUMENT, "workspace of %lld bytes; need %lld",
                    (long long)workspace_bytes, (long long)need);
    if ((n_lwes > 0 && (!masks || !bodies)) || !ksk_fft || !workspace || !out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const int N = (int)degree;
    int64_t *rows = reinterpret_cast<int64_t *>(workspace);
    cplx *partial = reinterpret_cast<cplx *>(rows + packs * n_in * degree);
    const int nb = packing_blocks(n_in);
    const int rpb = (int)((n_in + nb - 1) / nb);
    hipStream_t s = (hipStream_t)stream;
    if (n_lwes > 0) {
        dim3 tg((unsigned)((n_in + 63) / 64), (unsigned)((n_lwes + 63) / 64), (unsigned)packs);
        hipLaunchKernelGGL(pack_transpose_kernel, tg, dim3(256), 0, s, masks, (int)n_lwes, (int)n_in,
                           N, (int)log2_q, rows);
        if (int st = launch_status("packing transpose"))
            return st;
    }
    const cplx *kf = reinterpret_cast<const cplx *>(ksk_fft);
    BSLSTM_DEGREE_SWITCH(degree, hipLaunchKernelGGL((packing_mac_kernel<NN>), dim3(nb, (unsigned)packs),
                                                    dim3(NN / 8), 0, s, rows, (int)n_lwes, (int)n_in, rpb, kf,
                                                    (int)l, (int)log2_beta, (int)log2_q, partial,
                                                    slots, (int)slot_count));
    if (int st = launch_status("packing MAC"))
        return st;
    // the per-block partials summed across the chip into the (no longer
    // needed) transpose area: packs x 2 x N/2 complex fit there whenever
    // there is more than one block (n_in >= 2)
    const cplx *spec = partial;
    int spec_blocks = nb;
    if (nb > 1) {
        cplx *reduced = reinterpret_cast<cplx *>(rows);
        const int M = N / 2;
        hipLaunchKernelGGL(packing_reduce_kernel, dim3((unsigned)((M + RED_B - 1) / RED_B), 2, (unsigned)packs),
                           dim3(RED_G * RED_B), 0, s, partial, nb, M, reduced);
        if (int st = launch_status("packing reduce"))
            return st;
        spec = reduced;
        spec_blocks = 1;
    }
    BSLSTM_DEGREE_SWITCH(degree, hipLaunchKernelGGL((packing_finalize_kernel<NN>), dim3((unsigned)packs),
                                                    dim3(NN / 8), 0, s, spec, spec_blocks, bodies,
                                                    (int)n_lwes, (int)log2_q, out));
    return launch_status("packing finalize");
}

extern "C" int bslstm_packing(void *stream, const int64_t *masks, const int64_t *bodies,
                              int64_t packs, int64_t n_lwes, int64_t n_in, const double *ksk_fft,
                              int64_t degree, int64_t log2_beta, int64_t l, int64_t log2_q,
                              void *workspace, int64_t workspace_bytes, int64_t *out)
{
    return packing_run(stream, masks, bodies, packs, n_lwes, n_in, ksk_fft, nullptr, 1, degree,
                       log2_beta, l, log2_q, workspace, workspace_bytes, out);
}

extern "C" int bslstm_packing_slots(void *stream, const int64_t *masks, const int64_t *bodies,
                                    int64_t packs, int64_t n_lwes, int64_t n_in,
                                    const double *ksk_pool, const int32_t *slots,
                                    int64_t slot_count, int64_t degree, int64_t log2_beta,
                                    int64_t l, int64_t log2_q, void *workspace,
                                    int64_t workspace_bytes, int64_t *out)
{
    if (!slots || slot_count < 1 || slot_count > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "slots and 1 <= slot_count required");
    return packing_run(stream, masks, bodies, packs, n_lwes, n_in, ksk_pool, slots, slot_count,
                       degree, log2_beta, l, log2_q, workspace, workspace_bytes, out);
}

extern "C" int bslstm_torus_axpy(void *stream, const int64_t *x, const int64_t *y, int64_t n,
                                 int64_t sx, int64_t sy, int64_t log2_q, int64_t *out)
{
    clear_error();
    if (log2_q < 2 || log2_q > 64 || n < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad log2_q/n");
    if (n == 0)
        return BSLSTM_OK;
    if (!x || !y || !out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const int threads = 256;
    const int64_t blocks = std::min<int64_t>((n + threads - 1) / threads, 1 << 16);
    hipLaunchKernelGGL(torus_axpy_kernel, dim3((unsigned)blocks), dim3(threads), 0,
                       (hipStream_t)stream, x, y, n, sx, sy, (int)log2_q, out);
    return launch_status("torus axpy");
}

extern "C" int bslstm_rotate_poly(void *stream, const int64_t *in, int64_t rows, int64_t degree,
                                  const int64_t *exponents, int64_t rows_per_exponent,
                                  int64_t log2_q, int64_t *out)
{
    clear_error();
    if (!is_pow2(degree) || degree < 2 || degree > (1 << 20))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (log2_q < 2 || log2_q > 64 || rows < 0 || rows > 65535 || rows_per_exponent < 1)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad log2_q/rows (<= 65535)/rows_per_exponent");
    if (rows == 0)
        return BSLSTM_OK;
    if (!in || !exponents || !out || in == out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL or aliased operand");
    const int threads = 256;
    const int bx = (int)std::min<int64_t>((degree + threads - 1) / threads, 64);
    hipLaunchKernelGGL(rotate_poly_kernel, dim3(bx, (unsigned)rows), dim3(threads), 0,
                       (hipStream_t)stream, in, rows, (int)degree, exponents, rows_per_exponent,
                       (int)log2_q, out);
    return launch_status("polynomial rotation");
}
