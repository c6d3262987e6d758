// gfx950 kernels around the blind rotation: rotation exponents, the many-LUT
// rotate + sample-extract tail, and the LWE key switch.  In the reference
// these are XLA fusions generated from JAX code; here each is one
// hand-written, coalesced launch.
#include <algorithm>

#include "capi_internal.h"
#include "torus_device.h"

namespace bslstm {

// ---------------------------------------------------------------------------
// rotation exponents (cuTFHE/multiply.py:67-79)
// ---------------------------------------------------------------------------
__global__ void rotation_indices_kernel(const double *__restrict__ pk, int64_t total, int n,
                                        int c, int m, int n_iter, double two_n, double q_nlut,
                                        double n_lut, int two_N, int32_t *__restrict__ out)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total)
        return;
    const int p = (int)(gid % m);
    const int64_t bg = gid / m;
    const int g = (int)(bg % n_iter);
    const int64_t b = bg / n_iter;
    const double *row = pk + b * n + (int64_t)g * c;
    // jnp.dot(pk (..., n_iter, c), kron.T): bit j of pattern p is MSB-first
    double val = 0.0;
    for (int j = 0; j < c; j++)
        val += row[j] * (double)((p >> (c - 1 - j)) & 1);
    const double tq = (two_n * val) / q_nlut;
    const int32_t v = (int32_t)(rint(tq) * n_lut);
    int32_t r = v % two_N;
    if (r < 0)
        r += two_N;
    out[gid] = r;
}

// ---------------------------------------------------------------------------
// many-LUT tail: X^(-b_ms) rotation fused with sample extraction
// (schemas/RLWE_jax.py:295-303, text_jax.py:191-208, polynomial_jax.py:150-169)
// ---------------------------------------------------------------------------
__global__ void rotate_extract_kernel(const int64_t *__restrict__ br, const double *__restrict__ body,
                                      int N, int n_lut, double two_n, double q_nlut, double nl,
                                      double *__restrict__ masks, double *__restrict__ bodies)
{
    const int i = blockIdx.y;      // ciphertext
    const int k = blockIdx.z;      // extracted coefficient / sub-LUT
    const int64_t *A = br + (size_t)i * 2 * N;
    const int64_t *Bp = A + N;
    // alpha = -b_ms; the rotated polynomial R[s] = P[w] if w < N else -P[w-N],
    // w = (s - alpha) mod 2N
    const double bms = rint((two_n * body[i]) / q_nlut) * nl;
    const int alpha = -(int)bms;
    const int twoN = 2 * N;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
        // sample extract: a'_j = R_A[k-j] (j <= k) else -R_A[k-j+N]
        int s = k - j;
        int sign = 1;
        if (s < 0) {
            s += N;
            sign = -1;
        }
        int w = (s - alpha) % twoN;
        if (w < 0)
            w += twoN;
        int64_t v;
        if (w < N) {
            v = A[w];
        } else {
            v = -A[w - N];
        }
        masks[((size_t)i * n_lut + k) * N + j] = (double)(sign * v);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int w = (k - alpha) % twoN;
        if (w < 0)
            w += twoN;
        const int64_t v = w < N ? Bp[w] : -Bp[w - N];
        bodies[(size_t)i * n_lut + k] = (double)v;
    }
}

// ---------------------------------------------------------------------------
// LWE key switch (schemas/LWE_jax.py:97-117, gadget_product_LWE :65-82)
//
// out[r, j] = sum_{lev,i} d[r,lev,i] * KSK[lev,i,j], j in [0, n_out]; column
// n_out is the body key.  The digits are the reference's float rule on the
// input as given (x = a/q, no centring: polynomial_jax.py:90-111).  FP64
// products and sums are exact for torus32 (|sum| < 2^53), as in the
// reference's float64 tensordot, so partial sums over any split of (lev, i)
// add up exactly in any order: each workgroup owns RB ciphertexts x 256
// output columns x one range of input coefficients (the K split that fills
// the device), stages its digits in LDS, streams the key coalesced along j
// and adds its partial sums into the zeroed outputs with FP64 atomics; a
// final pass centres them (key_switch_final_kernel).  Deterministic: every
// partial sum is an integer below 2^53.
// ---------------------------------------------------------------------------
constexpr int KS_RB = 32;          // ciphertexts per workgroup
constexpr int KS_LDS_DOUBLES = 8192; // digit tile: l_ks * KS_RB * ic <= 64 KiB
constexpr int KS_LMAX = 64;        // max decomposition levels
constexpr int KS_THREADS = 256;
constexpr int KS_UNROLL = 4;       // key loads in flight per thread

__global__ void ks_zero_kernel(double *__restrict__ a_out, double *__restrict__ b_out, int64_t na,
                               int64_t nb)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < na + nb;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < na)
            a_out[e] = 0.0;
        else
            b_out[e - na] = 0.0;
    }
}

__global__ __launch_bounds__(KS_THREADS) void key_switch_kernel(
    const double *__restrict__ a_in, const double *__restrict__ ksk_a,
    const double *__restrict__ ksk_b, int batch, int n_in, int n_out, double beta, int l_ks, int ic_max,
    int i_per_split, double q, double *__restrict__ a_out, double *__restrict__ b_out)
{
    extern __shared__ double dig[];   // [l_ks][KS_RB][ic_max]
    const int j = blockIdx.x * KS_THREADS + threadIdx.x;
    const int r0 = blockIdx.y * KS_RB;
    const int i_begin = blockIdx.z * i_per_split;
    const int i_end = min(n_in, i_begin + i_per_split);
    const bool active = j <= n_out;
    double acc[KS_RB];
#pragma unroll
    for (int r = 0; r < KS_RB; r++)
        acc[r] = 0.0;

    for (int i0 = i_begin; i0 < i_end; i0 += ic_max) {
        const int ic = min(ic_max, i_end - i0);
        __syncthreads();
        for (int e = threadIdx.x; e < KS_RB * ic_max; e += KS_THREADS) {
            const int r = e / ic_max, ii = e % ic_max;
            const int row = r0 + r, i = i0 + ii;
            double x = (row < batch && ii < ic) ? a_in[(size_t)row * n_in + i] / q : 0.0;
            for (int lev = 0; lev < l_ks; lev++) {
                const double xb = x * beta;
                const double d = rint(xb);
                x = xb - d;
                dig[(lev * KS_RB + r) * ic_max + ii] = d;
            }
        }
        __syncthreads();
        if (active) {
            const double *kcol = j < n_out ? ksk_a + j : ksk_b;
            const size_t kstride = j < n_out ? (size_t)n_out : 1;
            for (int lev = 0; lev < l_ks; lev++) {
                const double *krow = kcol + ((size_t)lev * n_in + i0) * kstride;
                const double *drow = dig + (lev * KS_RB) * ic_max;
                // KS_UNROLL independent key loads in flight; digits past ic
                // are zero in LDS, so the tail adds zeros (exact)
                for (int ii = 0; ii < ic; ii += KS_UNROLL) {
                    double kv[KS_UNROLL];
#pragma unroll
                    for (int u = 0; u < KS_UNROLL; u++)
                        kv[u] = ii + u < ic ? krow[(size_t)(ii + u) * kstride] : 0.0;
#pragma unroll
                    for (int u = 0; u < KS_UNROLL; u++)
#pragma unroll
                        for (int r = 0; r < KS_RB; r++)
                            acc[r] = fma(drow[r * ic_max + ii + u], kv[u], acc[r]);
                }
            }
        }
    }
    if (!active)
        return;
#pragma unroll
    for (int r = 0; r < KS_RB; r++) {
        const int row = r0 + r;
        if (row >= batch)
            break;
        if (j < n_out)
            atomicAdd(&a_out[(size_t)row * n_out + j], acc[r]);
        else
            atomicAdd(&b_out[row], acc[r]);
    }
}

// centred mod q of the exact sums; the body adds b first (LWE_jax.py:117)
__global__ void key_switch_final_kernel(const double *__restrict__ b_in, int batch, int n_out,
                                        double q, double inv_q, double *__restrict__ a_out,
                                        double *__restrict__ b_out)
{
    const int64_t na = (int64_t)batch * n_out;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < na + batch;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < na) {
            const double v = a_out[e];
            a_out[e] = v - q * rint(v * inv_q);
        } else {
            const int64_t row = e - na;
            const double v = b_out[row] - q * rint(b_out[row] * inv_q);
            const double bb = b_in[row] + v;
            b_out[row] = bb - q * rint(bb * inv_q);
        }
    }
}


// ---------------------------------------------------------------------------
// int64 forms for moduli above 2^53 (the reference's q = 2^64 LSTM
// parameters, parameters_crypto.py:6): float64 cannot hold those classes, so
// ciphertexts are int64 centred representatives and sums wrap mod 2^64.
// ---------------------------------------------------------------------------
// rotation exponents from int64 masks: the group sum is formed exactly
// (mod 2^64, then centred mod q), then the reference's FP64 rounding
// round(2N*val/(q*n_lut))*n_lut mod 2N (multiply.py:67-79).
__global__ void rotation_indices_i64_kernel(const int64_t *__restrict__ pk, int64_t total, int n,
                                            int c, int m, int n_iter, int lq, double two_n,
                                            double q_nlut, double n_lut, int two_N,
                                            int32_t *__restrict__ out)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total)
        return;
    const int p = (int)(gid % m);
    const int64_t bg = gid / m;
    const int g = (int)(bg % n_iter);
    const int64_t b = bg / n_iter;
    const int64_t *row = pk + b * n + (int64_t)g * c;
    uint64_t acc = 0;
    for (int j = 0; j < c; j++)
        if ((p >> (c - 1 - j)) & 1)
            acc += (uint64_t)row[j];
    const double val = (double)tmod(acc, lq);
    const int32_t v = (int32_t)(rint((two_n * val) / q_nlut) * n_lut);
    int32_t r = v % two_N;
    if (r < 0)
        r += two_N;
    out[gid] = r;
}

// X^(-b_ms) rotation (skipped when body is NULL) + extraction of coefficients
// 0..n_ext-1 (RLWE_jax.py:295-303, text_jax.py:191-208), int64 in and out
__global__ void rotate_extract_i64_kernel(const int64_t *__restrict__ br,
                                          const int64_t *__restrict__ body, int N, int n_ext,
                                          double two_n, double q_nlut, double nl, int lq,
                                          int64_t *__restrict__ masks, int64_t *__restrict__ bodies)
{
    const int i = blockIdx.y;
    const int k = blockIdx.z;
    const int64_t *A = br + (size_t)i * 2 * N;
    const int64_t *Bp = A + N;
    const int alpha = body ? -(int)(rint((two_n * (double)body[i]) / q_nlut) * nl) : 0;
    const int twoN = 2 * N;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
        int s = k - j;
        bool neg = false;
        if (s < 0) {
            s += N;
            neg = true;
        }
        int w = (s - alpha) % twoN;
        if (w < 0)
            w += twoN;
        uint64_t v = (uint64_t)(w < N ? A[w] : A[w - N]);
        if ((w >= N) != neg)
            v = 0ULL - v;
        masks[((size_t)i * n_ext + k) * N + j] = tmod(v, lq);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int w = (k - alpha) % twoN;
        if (w < 0)
            w += twoN;
        const uint64_t v = w < N ? (uint64_t)Bp[w] : 0ULL - (uint64_t)Bp[w - N];
        bodies[(size_t)i * n_ext + k] = tmod(v, lq);
    }
}

// exact int64 key switch: a' = sum d*KSK_a, b' = b + sum d*KSK_b mod 2^log2_q,
// digits by the float rule restated on integers (polynomial_jax.py:90-111).
// The reduction over the n_in inputs is split across gridDim.z workgroups
// (chunks of ic_max inputs, round robin): at LSTM batch sizes (tens of
// ciphertexts) the output tile alone gives only a handful of workgroups.
// Partial sums meet in the outputs through 64-bit integer atomics, which
// are exact and order-independent, so the result is bit-reproducible.
// key_switch_i64_init writes a' = 0, b' = b first; key_switch_i64_final
// reduces to the centred representative afterwards.
__global__ void key_switch_i64_init(const int64_t *__restrict__ b_in, int batch, int n_out,
                                    int64_t *__restrict__ a_out, int64_t *__restrict__ b_out)
{
    const int64_t total = (int64_t)batch * (n_out + 1);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / (n_out + 1);
        const int j = (int)(e % (n_out + 1));
        if (j < n_out)
            a_out[row * n_out + j] = 0;
        else
            b_out[row] = b_in[row];
    }
}

__global__ void key_switch_i64_final(int batch, int n_out, int lq, int64_t *__restrict__ a_out,
                                     int64_t *__restrict__ b_out)
{
    const int64_t total = (int64_t)batch * (n_out + 1);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / (n_out + 1);
        const int j = (int)(e % (n_out + 1));
        int64_t *p = j < n_out ? a_out + row * n_out + j : b_out + row;
        *p = tmod((uint64_t)*p, lq);
    }
}

__global__ __launch_bounds__(KS_THREADS) void key_switch_i64_kernel(
    const int64_t *__restrict__ a_in, const int64_t *__restrict__ ksk_a,
    const int64_t *__restrict__ ksk_b, int batch, int n_in, int n_out, int lb, int l_ks,
    int ic_max, int lq, int64_t *__restrict__ a_out, int64_t *__restrict__ b_out)
{
    extern __shared__ int32_t digi[];   // [l_ks][KS_RB][ic_max]
    const int j = blockIdx.x * KS_THREADS + threadIdx.x;
    const int r0 = blockIdx.y * KS_RB;
    const bool active = j <= n_out;
    uint64_t acc[KS_RB];
#pragma unroll
    for (int r = 0; r < KS_RB; r++)
        acc[r] = 0;
    for (int i0 = blockIdx.z * ic_max; i0 < n_in; i0 += gridDim.z * ic_max) {
        __syncthreads();
        for (int e = threadIdx.x; e < KS_RB * ic_max; e += KS_THREADS) {
            const int r = e / ic_max, ii = e % ic_max;
            const int row = r0 + r, i = i0 + ii;
            int64_t x = (row < batch && i < n_in) ? tmod((uint64_t)a_in[(size_t)row * n_in + i], lq) : 0;
            for (int lev = 0; lev < l_ks; lev++) {
                const int sh = lq - (lev + 1) * lb;
                int64_t d;
                if (sh <= 0) {
                    d = x;
                    x = 0;
                } else {
                    d = rhe_shift(x, sh);
                    x = (int64_t)((uint64_t)x - ((uint64_t)d << sh));
                }
                digi[(lev * KS_RB + r) * ic_max + ii] = (int32_t)d;
            }
        }
        __syncthreads();
        if (active) {
            const int ic = min(ic_max, n_in - i0);
            for (int lev = 0; lev < l_ks; lev++) {
                for (int ii = 0; ii < ic; ii++) {
                    const size_t kidx = (size_t)lev * n_in + i0 + ii;
                    const uint64_t kv = (uint64_t)(j < n_out ? ksk_a[kidx * n_out + j] : ksk_b[kidx]);
                    const int32_t *dr = digi + (lev * KS_RB) * ic_max + ii;
#pragma unroll
                    for (int r = 0; r < KS_RB; r++)
                        acc[r] += (uint64_t)(int64_t)dr[r * ic_max] * kv;
                }
            }
        }
    }
    if (!active)
        return;
#pragma unroll
    for (int r = 0; r < KS_RB; r++) {
        const int row = r0 + r;
        if (row >= batch)
            break;
        unsigned long long *dst = reinterpret_cast<unsigned long long *>(
            j < n_out ? a_out + (size_t)row * n_out + j : b_out + row);
        atomicAdd(dst, (unsigned long long)acc[r]);
    }
}

}  // namespace bslstm

using namespace bslstm;

extern "C" int bslstm_rotation_indices(void *stream, const double *public_key, int64_t batch,
                                       int64_t n, int64_t collapse, int64_t log2_q, int64_t degree,
                                       int64_t n_lut, int32_t *indices)
{
    clear_error();
    if (collapse < 1 || collapse > 6 || n % collapse != 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "n=%lld must be a multiple of collapse=%lld in [1,6]",
                    (long long)n, (long long)collapse);
    if (!is_pow2(degree) || degree > (1 << 20))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (n_lut < 1 || log2_q < 1 || log2_q > 64 || batch < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad n_lut/log2_q/batch");
    const int n_iter = (int)(n / collapse), m = 1 << collapse;
    const int64_t total = batch * n_iter * m;
    if (total == 0)
        return BSLSTM_OK;
    if (!public_key || !indices)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const double q = ldexp(1.0, (int)log2_q);
    const int threads = 256;
    const int64_t blocks = (total + threads - 1) / threads;
    hipLaunchKernelGGL(rotation_indices_kernel, dim3((unsigned)blocks), dim3(threads), 0,
                       (hipStream_t)stream, public_key, total, (int)n, (int)collapse, m, n_iter,
                       (double)(2 * degree), q * (double)n_lut, (double)n_lut, (int)(2 * degree),
                       indices);
    return launch_status("rotation indices");
}

extern "C" int bslstm_rotate_extract(void *stream, const int64_t *br_out, const double *body,
                                     int64_t batch, int64_t degree, int64_t log2_q, int64_t n_lut,
                                     double *masks, double *bodies)
{
    clear_error();
    if (!is_pow2(degree) || degree > 65536)
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (n_lut < 1 || n_lut > degree || log2_q < 1 || log2_q > 64 || batch < 0 || batch > 65535)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad n_lut/log2_q/batch (batch <= 65535 per call)");
    if (batch == 0)
        return BSLSTM_OK;
    if (!br_out || !body || !masks || !bodies)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const double q = ldexp(1.0, (int)log2_q);
    const int threads = 256;
    const int bx = (int)((degree + threads - 1) / threads);
    hipLaunchKernelGGL(rotate_extract_kernel, dim3(bx, (unsigned)batch, (unsigned)n_lut),
                       dim3(threads), 0, (hipStream_t)stream, br_out, body, (int)degree, (int)n_lut,
                       (double)(2 * degree), q * (double)n_lut, (double)n_lut, masks, bodies);
    return launch_status("rotate/extract");
}

extern "C" int bslstm_key_switch_lwe(void *stream, const double *a_in, const double *b_in,
                                     const double *ksk_a, const double *ksk_b, int64_t batch,
                                     int64_t n_in, int64_t n_out, int64_t log2_beta_ks, int64_t l_ks,
                                     int64_t log2_q, double *a_out, double *b_out)
{
    clear_error();
    if (log2_q < 1 || log2_q > 32)
        return fail(BSLSTM_INVALID_ARGUMENT, "key switch needs log2_q <= 32 (exact FP64 sums)");
    if (l_ks < 1 || l_ks > KS_LMAX || log2_beta_ks < 1 || l_ks * log2_beta_ks > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad l_ks/log2_beta_ks");
    if (n_in < 1 || n_out < 1 || batch < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad shapes");
    // exactness bound: l*n_in*(beta/2)*(q/2) < 2^53
    const double bound = (double)l_ks * (double)n_in * ldexp(1.0, (int)log2_beta_ks - 1) *
                         ldexp(1.0, (int)log2_q - 1);
    if (bound >= 0x1p53)
        return fail(BSLSTM_INVALID_ARGUMENT, "key-switch sums could exceed 2^53");
    if (batch == 0)
        return BSLSTM_OK;
    if (!a_in || !b_in || !ksk_a || !ksk_b || !a_out || !b_out)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    // the outputs are zeroed first and accumulated into: they must not
    // overlap the inputs
    {
        auto overlap = [](const void *p, int64_t np, const void *o, int64_t no) {
            const char *a = (const char *)p, *b = (const char *)o;
            return a < b + no && b < a + np;
        };
        const int64_t bytes_a_out = batch * n_out * 8, bytes_b = batch * 8;
        if (overlap(a_in, batch * n_in * 8, a_out, bytes_a_out) ||
            overlap(a_in, batch * n_in * 8, b_out, bytes_b) ||
            overlap(b_in, bytes_b, a_out, bytes_a_out) || overlap(b_in, bytes_b, b_out, bytes_b))
            return fail(BSLSTM_INVALID_ARGUMENT, "key-switch outputs overlap its inputs");
    }
    const double q = ldexp(1.0, (int)log2_q);
    hipStream_t s = (hipStream_t)stream;
    const int64_t na = batch * n_out;
    hipLaunchKernelGGL(ks_zero_kernel, dim3((unsigned)std::min<int64_t>(1024, (na + batch + 255) / 256)),
                       dim3(256), 0, s, a_out, b_out, na, batch);
    const unsigned col_tiles = (unsigned)((n_out + 1 + KS_THREADS - 1) / KS_THREADS);
    const unsigned row_tiles = (unsigned)((batch + KS_RB - 1) / KS_RB);
    // staging width: a multiple of the unroll (the digits past a short last
    // chunk are zero in LDS, and every read stays inside the tile); >= 4
    // since l_ks <= KS_LMAX
    const int ic_max = (int)(std::min<int64_t>(32, KS_LDS_DOUBLES / (l_ks * KS_RB)) / KS_UNROLL *
                             KS_UNROLL);
    // split the input coefficients so that the launch has ~2048 workgroups
    // (8 per CU), each range a multiple of the staging width
    const int64_t tiles = (int64_t)col_tiles * row_tiles;
    const int64_t max_splits = (n_in + ic_max - 1) / ic_max;
    const int64_t splits = std::max<int64_t>(1, std::min<int64_t>(max_splits, (2048 + tiles - 1) / tiles));
    const int64_t chunks = (max_splits + splits - 1) / splits;
    const int i_per_split = (int)(chunks * ic_max);
    const unsigned zsplits = (unsigned)((n_in + i_per_split - 1) / i_per_split);
    const size_t lds = sizeof(double) * (size_t)l_ks * KS_RB * ic_max;
    hipLaunchKernelGGL(key_switch_kernel, dim3(col_tiles, row_tiles, zsplits), dim3(KS_THREADS), lds,
                       s, a_in, ksk_a, ksk_b, (int)batch, (int)n_in, (int)n_out,
                       ldexp(1.0, (int)log2_beta_ks), (int)l_ks, ic_max, i_per_split, q, a_out, b_out);
    hipLaunchKernelGGL(key_switch_final_kernel, dim3((unsigned)std::min<int64_t>(1024, (na + batch + 255) / 256)),
                       dim3(256), 0, s, b_in, (int)batch, (int)n_out, q, 1.0 / q, a_out, b_out);
    return launch_status("key switch");
}

extern "C" int bslstm_rotation_indices_i64(void *stream, const int64_t *public_key, int64_t batch,
                                           int64_t n, int64_t collapse, int64_t log2_q,
                                           int64_t degree, int64_t n_lut, int32_t *indices)
{
    clear_error();
    if (collapse < 1 || collapse > 6 || n % collapse != 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "n=%lld must be a multiple of collapse=%lld in [1,6]",
                    (long long)n, (long long)collapse);
    if (!is_pow2(degree) || degree > (1 << 20))
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (n_lut < 1 || log2_q < 2 || log2_q > 64 || batch < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad n_lut/log2_q/batch");
    const int n_iter = (int)(n / collapse), m = 1 << collapse;
    const int64_t total = batch * n_iter * m;
    if (total == 0)
        return BSLSTM_OK;
    if (!public_key || !indices)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const double q = ldexp(1.0, (int)log2_q);
    const int threads = 256;
    const int64_t blocks = (total + threads - 1) / threads;
    hipLaunchKernelGGL(rotation_indices_i64_kernel, dim3((unsigned)blocks), dim3(threads), 0,
                       (hipStream_t)stream, public_key, total, (int)n, (int)collapse, m, n_iter,
                       (int)log2_q, (double)(2 * degree), q * (double)n_lut, (double)n_lut,
                       (int)(2 * degree), indices);
    return launch_status("rotation indices");
}

extern "C" int bslstm_rotate_extract_i64(void *stream, const int64_t *br_out, const int64_t *body,
                                         int64_t batch, int64_t degree, int64_t log2_q,
                                         int64_t n_lut, int64_t n_ext, int64_t *masks,
                                         int64_t *bodies)
{
    clear_error();
    if (!is_pow2(degree) || degree > 65536)
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    if (n_lut < 1 || n_ext < 1 || n_ext > degree || n_ext > 65535 || log2_q < 2 || log2_q > 64 ||
        batch < 0 || batch > 65535)
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "bad n_lut/n_ext/log2_q/batch (n_ext <= degree, batch <= 65535 per call)");
    if (batch == 0)
        return BSLSTM_OK;
    if (!br_out || !masks || !bodies)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const double q = ldexp(1.0, (int)log2_q);
    const int threads = 256;
    const int bx = (int)((degree + threads - 1) / threads);
    hipLaunchKernelGGL(rotate_extract_i64_kernel, dim3(bx, (unsigned)batch, (unsigned)n_ext),
                       dim3(threads), 0, (hipStream_t)stream, br_out, body, (int)degree,
                       (int)n_ext, (double)(2 * degree), q * (double)n_lut, (double)n_lut,
                       (int)log2_q, masks, bodies);
    return launch_status("rotate/extract");
}

extern "C" int bslstm_key_switch_lwe_i64(void *stream, const int64_t *a_in, const int64_t *b_in,
                                         const int64_t *ksk_a, const int64_t *ksk_b, int64_t batch,
                                         int64_t n_in, int64_t n_out, int64_t log2_beta_ks,
                                         int64_t l_ks, int64_t log2_q, int64_t *a_out,
                                         int64_t *b_out)
{
    clear_error();
    if (log2_q < 2 || log2_q > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "log2_q must be in [2, 64]");
    if (l_ks < 1 || l_ks > KS_LMAX || log2_beta_ks < 1 || log2_beta_ks > 30 ||
        l_ks * log2_beta_ks > log2_q)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad l_ks/log2_beta_ks");
    if (n_in < 1 || n_out < 1 || batch < 0 || batch > (1LL << 30))
        return fail(BSLSTM_INVALID_ARGUMENT, "bad shapes");
    if (batch == 0)
        return BSLSTM_OK;
    if (!a_in || !b_in || !ksk_a || !ksk_b || !a_out || !b_out || b_out == b_in || a_out == a_in)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL or aliased operand");
    hipStream_t s = (hipStream_t)stream;
    const int64_t total = batch * (n_out + 1);
    const unsigned eb = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(key_switch_i64_init, dim3(eb), dim3(256), 0, s, b_in, (int)batch,
                       (int)n_out, a_out, b_out);
    const int ic_max = (int)std::min<int64_t>(32, (2 * KS_LDS_DOUBLES) / (l_ks * KS_RB));
    const int64_t gx = (n_out + 1 + KS_THREADS - 1) / KS_THREADS, gy = (batch + KS_RB - 1) / KS_RB;
    const int64_t chunks = (n_in + ic_max - 1) / ic_max;
    // enough workgroups to fill the chip (~2 per CU) when the tile grid is small
    const int64_t gz = std::max<int64_t>(1, std::min<int64_t>(chunks, 512 / (gx * gy)));
    const size_t lds = sizeof(int32_t) * (size_t)l_ks * KS_RB * ic_max;
    hipLaunchKernelGGL(key_switch_i64_kernel, dim3((unsigned)gx, (unsigned)gy, (unsigned)gz),
                       dim3(KS_THREADS), lds, s, a_in, ksk_a, ksk_b, (int)batch, (int)n_in,
                       (int)n_out, (int)log2_beta_ks, (int)l_ks, ic_max, (int)log2_q, a_out, b_out);
    if (int st = launch_status("key switch"))
        return st;
    if (log2_q < 64)
        hipLaunchKernelGGL(key_switch_i64_final, dim3(eb), dim3(256), 0, s, (int)batch, (int)n_out,
                           (int)log2_q, a_out, b_out);
    return launch_status("key switch");
}
