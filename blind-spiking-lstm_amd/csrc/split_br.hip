// Collapsed blind rotation in exact integer arithmetic on the matrix cores
// (the "split" form): int8 MFMA GEMMs with int32 accumulation, the
// bootstrapping key cut into signed byte limbs.  Same semantics as the FP64
// forms (csrc/mul_bootstrap.cuh:20-136 of the reference): per step g
//   dec[bc]  = offset-rule digits of the accumulator (csrc/dec_fft.cuh:4-31)
//   ct[a]    = sum_{bc,p} BSK[g,a,bc,p] * X^idx[g,p] * dec[bc]   mod q
// but every product is exact, so the result is the one the FP64 transform
// rounds towards; it equals the int128 oracle bit for bit.  log2_q <= 32.
//
// Method (one Nussbaumer level, no roots of unity, no modular reduction):
//   N = M*R with R | M.  a(X) = sum_{u<R} X^u A_u(Y), Y = X^R, A_u in
//   Z[Y]/(Y^M+1) with A_u[v] = a[u + R v].  w = Y^(M/R) has order 2R, so a
//   length-2R DFT over that ring moves and negates coefficients only.  The
//   2R bins of a product are negacyclic products of size M; the inverse
//   (without 1/2R) gives 2R times the length-2R cyclic convolution, which
//   folds (X^R = Y) into 2R times the product.  Everything is Z-linear with
//   integer coefficients, so it runs mod 2^64 and the exact division by 2R
//   leaves the product mod 2^32.
//
// Per step and ciphertext the 2l*m rotated digit polynomials X^e_p dec_bc
// (digits |d| <= beta/2 <= 64) are transformed: |bin coefficients| <= R*64,
// two signed byte limbs.  The key's bins (mod 2^(32+log2 2R)) are cut into
// five signed byte limbs once.  Per bin, the sum over (p, bc) of size-M
// negacyclic products is one GEMM
//   C[(a,k), ct] = sum_{(pbc,i)} T_{pbc,a}[k,i] * D[ct][pbc][i]
// whose A operand is the key's negacyclic Toeplitz matrix (shared by every
// ciphertext: the MFMA's reuse) and whose B operand holds the ciphertexts'
// transformed digits.  Limb pairs (key t, digit s) with the same weight
// 2^(8(t+s)) share one int32 accumulator (|sum| < 2^31); pairs of weight
// >= 2^(32+log2 2R) are dropped (they vanish mod that power).
//
// Kernels per step (batch processed in chunks that fit the workspace):
//   split_step_kernel  inverse transform of the previous step's bins, fold,
//                      division, digits, rotations, forward transforms of the
//                      rotated digits -> int8 limb tiles (B operand)
//   split_gemm_kernel  one workgroup per (bin, group of ciphertexts): MFMA
//                      over K = (pbc, i) with the Toeplitz rows built from
//                      the key's bin in LDS -> int64 bins of the products
// The transforms run in registers, one ring element coefficient per lane
// (lane v holds coefficient v of every element), so the multiplications by
// powers of Y are uniform cross-lane shifts (shuffles) with a sign.
#include <algorithm>

#include "capi_internal.h"
#include "tfhe_device.h"

namespace bslstm {

typedef int v4i_s __attribute__((ext_vector_type(4)));
typedef int v16i_s __attribute__((ext_vector_type(16)));

constexpr int SPLIT_LIMBS = 5;      // key limbs (40 bits >= 32 + log2(2R))
constexpr int SPLIT_PC = 16;        // pbc per key chunk in the GEMM (LDS)
constexpr int SPLIT_THREADS = 256;

template <int N> struct Ring;
template <> struct Ring<512> { static constexpr int M = 32, R = 16; };
template <> struct Ring<1024> { static constexpr int M = 32, R = 32; };
template <> struct Ring<2048> { static constexpr int M = 64, R = 32; };

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// x * Y^S in Z[Y]/(Y^M+1) for the ring element whose coefficient v sits in
// lane base + v (M = 32: two elements per wave, base 0 or 32).  S is a
// compile-time shift in [0, 2M).
template <int M, int S, typename T>
__device__ __forceinline__ T mul_y(T x, int v, int base)
{
    constexpr int s = S % (2 * M);
    constexpr int sm = s % M;
    constexpr bool neg_all = s >= M;
    if constexpr (sm == 0) {
        return neg_all ? (T)(-x) : x;
    } else {
        T y = __shfl(x, base | ((v - sm) & (M - 1)));
        const bool neg = (v < sm) != neg_all;
        return neg ? (T)(-y) : y;
    }
}

// Forward transform (decimation in frequency, bit-reversed bin order) of the
// zero-padded sequence x[0 .. R-1] (x[R .. 2R-1] are outputs only).
template <int M, int R, typename T>
__device__ __forceinline__ void nuss_forward(T (&x)[2 * R], int v, int base)
{
    constexpr int L = 2 * R;
    // first stage: the upper half of the input is zero
    [&]<int... K>(std::integer_sequence<int, K...>) {
        ((x[K + R] = mul_y<M, K * (M / R)>(x[K], v, base)), ...);
    }(std::make_integer_sequence<int, R>{});
    [&]<int... H>(std::integer_sequence<int, H...>) {
        (([&] {
             constexpr int half = R >> (H + 1);
             constexpr int step = (L / (2 * half)) * (M / R);
             [&]<int... J>(std::integer_sequence<int, J...>) {
                 (([&] {
                      constexpr int b = (J / half) * 2 * half, k = J % half;
                      const T p = x[b + k], q = x[b + k + half];
                      x[b + k] = p + q;
                      x[b + k + half] = mul_y<M, k * step>((T)(p - q), v, base);
                  }()), ...);
             }(std::make_integer_sequence<int, R>{});
         }()), ...);
    }(std::make_integer_sequence<int, ilog2(R)>{});
}

// Inverse (decimation in time on bit-reversed input, w^-1, no 1/2R).
template <int M, int R, typename T>
__device__ __forceinline__ void nuss_inverse(T (&x)[2 * R], int v, int base)
{
    constexpr int L = 2 * R;
    [&]<int... H>(std::integer_sequence<int, H...>) {
        (([&] {
             constexpr int half = 1 << H;
             constexpr int step = (L / (2 * half)) * (M / R);
             [&]<int... J>(std::integer_sequence<int, J...>) {
                 (([&] {
                      constexpr int b = (J / half) * 2 * half, k = J % half;
                      const T p = x[b + k];
                      const T q = mul_y<M, (2 * M - (k * step) % (2 * M)) % (2 * M)>(x[b + k + half], v, base);
                      x[b + k] = p + q;
                      x[b + k + half] = p - q;
                  }()), ...);
             }(std::make_integer_sequence<int, R>{});
         }()), ...);
    }(std::make_integer_sequence<int, ilog2(L)>{});
}

// Balanced signed byte limb t of v mod 2^40 (sum_t limb_t 2^(8t) = v mod 2^40).
__device__ __forceinline__ int byte_limb(uint64_t v, int t)
{
    int carry = 0, d = 0;
    for (int i = 0; i <= t; i++) {
        d = (int)((v >> (8 * i)) & 255) + carry;
        carry = d >= 128;
        d -= carry * 256;
    }
    return d;
}

// ---------------------------------------------------------------------------
// Key preparation: one workgroup per key polynomial (g, a, lev, comp, p) of
// the FFT-domain key (n_iter, 2, l, 2, m, N/2) c128.  Inverse transform
// (the reference convention F_j = p(zeta^(2j+1)), polynomial_jax.py:17-21),
// rounding to the integer coefficients, Nussbaumer forward in int64, then
// per bin the Toeplitz source rows
//   rev[x] = ext[M-1-x], ext[d] = K[d] (d >= 0), -K[d+M] (d < 0), rev[2M-1] = 0
// as five byte-limb rows, layout key[g][bin][pbc][a][t][2M], pbc = bc*m + p.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(SPLIT_THREADS) void split_key_kernel(const cplx *__restrict__ bsk,
                                                                  int L, int m, int P_pad,
                                                                  int8_t *__restrict__ key)
{
    constexpr int M = Ring<N>::M, R = Ring<N>::R, NB = 2 * R;
    __shared__ cplx z[N];          // spectrum, then time domain
    __shared__ cplx tw[N / 2];     // exp(2 pi i k / N)
    __shared__ int64_t xs[2 * N];  // Nussbaumer elements [u][v]
    const int poly = blockIdx.x;   // ((g*2 + a)*L + lev)*2 + comp)*m + p
    const int p = poly % m;
    int rest = poly / m;
    const int comp = rest % 2;
    rest /= 2;
    const int lev = rest % L;
    rest /= L;
    const int a = rest % 2;
    const int g = rest / 2;
    const int pbc = (2 * lev + comp) * m + p;
    const cplx *F = bsk + (size_t)poly * (N / 2);
    for (int k = threadIdx.x; k < N / 2; k += blockDim.x) {
        double s, c;
        sincospi(2.0 * k / N, &s, &c);
        tw[k] = {c, s};
    }
    // Hermitian extension, written at bit-reversed positions
    constexpr int LG = ilog2(N);
    for (int j = threadIdx.x; j < N / 2; j += blockDim.x) {
        const cplx f = F[j];
        z[__brev((unsigned)j) >> (32 - LG)] = f;
        z[__brev((unsigned)(N - 1 - j)) >> (32 - LG)] = cconj(f);
    }
    __syncthreads();
    for (int half = 1; half < N; half *= 2) {
        for (int e = threadIdx.x; e < N / 2; e += blockDim.x) {
            const int k = e % half, b = (e / half) * 2 * half;
            const cplx w = tw[k * (N / (2 * half))];
            const cplx u = z[b + k], t = cmul(z[b + k + half], w);
            z[b + k] = cadd(u, t);
            z[b + k + half] = csub(u, t);
        }
        __syncthreads();
    }
    // p_k = Re(y_k * exp(i pi k / N)) / N, integer-valued
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        double s, c;
        sincospi((double)k / N, &s, &c);
        const double val = (z[k].re * c - z[k].im * s) / N;
        const int u = k % R, v = k / R;
        xs[u * M + v] = (int64_t)rint(val);
        xs[(u + R) * M + v] = 0;
    }
    __syncthreads();
    // Nussbaumer forward (decimation in frequency), element pairs per stage
    for (int half = R; half >= 1; half /= 2) {
        const int step = (2 * R / (2 * half)) * (M / R);
        int64_t r0[N / SPLIT_THREADS], r1[N / SPLIT_THREADS];
        int n = 0;
        for (int e = threadIdx.x; e < R * M; e += blockDim.x, n++) {
            const int v = e % M, J = e / M;
            const int k = J % half, b = (J / half) * 2 * half;
            const int s = (k * step) % (2 * M);
            const int sm = s % M;
            const int vs = (v - sm) & (M - 1);
            const bool neg = (v < sm) != (s >= M);
            const int64_t p0 = xs[(b + k) * M + v], q0 = xs[(b + k + half) * M + v];
            const int64_t d = xs[(b + k) * M + vs] - xs[(b + k + half) * M + vs];
            r0[n] = p0 + q0;
            r1[n] = neg ? -d : d;
        }
        __syncthreads();
        n = 0;
        for (int e = threadIdx.x; e < R * M; e += blockDim.x, n++) {
            const int v = e % M, J = e / M;
            const int k = J % half, b = (J / half) * 2 * half;
            xs[(b + k) * M + v] = r0[n];
            xs[(b + k + half) * M + v] = r1[n];
        }
        __syncthreads();
    }
    // Toeplitz source rows, four bytes per store
    for (int e = threadIdx.x; e < NB * SPLIT_LIMBS * (2 * M / 4); e += blockDim.x) {
        const int x4 = e % (2 * M / 4);
        const int t = (e / (2 * M / 4)) % SPLIT_LIMBS;
        const int bin = e / (2 * M / 4) / SPLIT_LIMBS;
        uint32_t word = 0;
        for (int j = 0; j < 4; j++) {
            const int x = 4 * x4 + j;
            int limb = 0;
            if (x < 2 * M - 1) {
                const int d = M - 1 - x;
                const int64_t val = d >= 0 ? xs[bin * M + d] : -xs[bin * M + d + M];
                limb = byte_limb((uint64_t)val, t);
            }
            word |= (uint32_t)(uint8_t)(int8_t)limb << (8 * j);
        }
        int8_t *dst = key + (((((size_t)g * NB + bin) * P_pad + pbc) * 2 + a) * SPLIT_LIMBS + t) * (2 * M);
        reinterpret_cast<uint32_t *>(dst)[x4] = word;
    }
}

// zero rows of the padded pbc range [P, P_pad)
__global__ void split_key_pad_kernel(int8_t *__restrict__ key, int64_t blocks, int P, int P_pad,
                                     int row_bytes)
{
    const int64_t words = (int64_t)(P_pad - P) * row_bytes / 4;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < blocks * words;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t blk = e / words, w = e % words;
        reinterpret_cast<uint32_t *>(key + (blk * P_pad + P) * (int64_t)row_bytes)[w] = 0;
    }
}

// ---------------------------------------------------------------------------
// Step kernel: one workgroup of four waves per ciphertext of the chunk.
// ---------------------------------------------------------------------------
struct SplitShape {
    int n_iter, m, P, P_pad, log2_beta, log2_q;
};

template <int N, int L>
__global__ __launch_bounds__(SPLIT_THREADS) void split_step_kernel(
    const int64_t *__restrict__ ct_init, int64_t ct_stride, const int64_t *__restrict__ bins,
    const int32_t *__restrict__ indices, int8_t *__restrict__ dtiles, int64_t *__restrict__ out,
    int count, int g, SplitShape sh)
{
    constexpr int M = Ring<N>::M, R = Ring<N>::R, NB = 2 * R, LG = ilog2(2 * R);
    constexpr int PER_WAVE = 64 / M;                 // ring elements per wave (1 or 2)
    const int ct = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int v = lane & (M - 1), base = lane & ~(M - 1);
    const int sub = lane / M;                        // which element of the wave
    __shared__ int8_t dig[2 * L][R][M];              // digits [bc][u][v] = dec_bc[u + R v]
    __shared__ __align__(16) int8_t stage[4][2][NB][M * PER_WAVE];
    const int lq = sh.log2_q;
    const uint32_t qmask = lq == 32 ? 0xFFFFFFFFu : ((1u << lq) - 1);

    // ---- accumulator rows (lane v holds coefficients u + R v, u < R)
    const int a = w * PER_WAVE + sub;                // row of this lane's element
    if (a < 2) {
        uint32_t c32[R];
        if (g == 0) {
            const int64_t *src = ct_init + (size_t)ct * ct_stride + (size_t)a * N;
#pragma unroll
            for (int u = 0; u < R; u++)
                c32[u] = (uint32_t)src[u + R * v];
        } else {
            int64_t x[2 * R];
            const int64_t *src = bins + ((size_t)ct * 2 + a) * NB * M;
#pragma unroll
            for (int j = 0; j < NB; j++)
                x[j] = src[j * M + v];
            nuss_inverse<M, R>(x, v, base);
#pragma unroll
            for (int u = 0; u < R; u++) {
                const int64_t cu = x[u] + mul_y<M, 1>(x[u + R], v, base);
                c32[u] = (uint32_t)((uint64_t)cu >> LG);
            }
        }
        if (g == sh.n_iter) {
            int64_t *dst = out + ((size_t)ct * 2 + a) * N;
            const int s = 64 - lq;
#pragma unroll
            for (int u = 0; u < R; u++)
                dst[u + R * v] = (int64_t)((uint64_t)(c32[u] & qmask) << s) >> s;
        } else {
            uint32_t off = 0;
            for (int i = 1; i <= L; i++)
                off += 1u << (lq - i * sh.log2_beta + sh.log2_beta - 1);
            if (lq - L * sh.log2_beta - 1 >= 0)
                off += 1u << (lq - L * sh.log2_beta - 1);
            const uint32_t dmask = (1u << sh.log2_beta) - 1;
            const int hb = 1 << (sh.log2_beta - 1);
#pragma unroll
            for (int u = 0; u < R; u++) {
                const uint32_t st = ((c32[u] & qmask) + off) & qmask;
#pragma unroll
                for (int lev = 0; lev < L; lev++)
                    dig[2 * lev + a][u][v] =
                        (int8_t)((int)((st >> (lq - (lev + 1) * sh.log2_beta)) & dmask) - hb);
            }
        }
    }
    if (g == sh.n_iter)
        return;
    __syncthreads();

    // ---- rotated digit polynomials, forward transforms, limb tiles
    const int32_t *idp = indices + ((size_t)ct * sh.n_iter + g) * sh.m;
    const int tile = ct >> 5, n = ct & 31;
    const int KC = sh.P_pad * (M / 16);
    int8_t *tbase = dtiles + (size_t)tile * NB * 2 * KC * 512;
    for (int pb0 = w * PER_WAVE; pb0 < sh.P_pad; pb0 += 4 * PER_WAVE) {
        const int pbc = pb0 + sub;
        int x[2 * R];
        if (pbc < sh.P) {
            const int bc = pbc / sh.m, p = pbc - bc * sh.m;
            const int e = idp[p] & (2 * N - 1);
            const bool flip = e >= N;
            const int e1 = e & (N - 1);
#pragma unroll
            for (int u = 0; u < R; u++) {
                int j = u + R * v - e1;       // coefficient of dec_bc feeding X^e dec_bc
                const bool neg = (j < 0) != flip;
                j += j < 0 ? N : 0;
                const int d = dig[bc][j % R][j / R];
                x[u] = neg ? -d : d;
            }
        } else {
#pragma unroll
            for (int u = 0; u < R; u++)
                x[u] = 0;
        }
        nuss_forward<M, R>(x, v, base);
        // limbs to LDS [limb][bin][element*M + v], then 16-byte chunks out
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const int lo = (int)(int8_t)(x[j] & 255);
            stage[w][0][j][sub * M + v] = (int8_t)lo;
            stage[w][1][j][sub * M + v] = (int8_t)((x[j] - lo) >> 8);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // chunks: (limb s, bin j, element e, 16-coefficient chunk c)
        constexpr int CH = 2 * NB * PER_WAVE * (M / 16);
        for (int q = lane; q < CH; q += 64) {
            const int c = q % (M / 16);
            const int el = (q / (M / 16)) % PER_WAVE;
            const int j = (q / (M / 16) / PER_WAVE) % NB;
            const int s = q / (M / 16) / PER_WAVE / NB;
            const int pe = pb0 + el;
            const v4i_s val = *reinterpret_cast<const v4i_s *>(&stage[w][s][j][el * M + c * 16]);
            const int kc = pe * (M / 16) + c;
            *reinterpret_cast<v4i_s *>(tbase + ((size_t)j * 2 + s) * KC * 512 +
                                       ((size_t)(kc >> 3) * 32 + n) * 128 + (kc & 7) * 16) = val;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// GEMM kernel: workgroup (bin, group of ciphertexts).  Rows (a, k), M per
// output row a: RT = 2M/32 row tiles.  M = 64: wave w takes row tile w and
// two column tiles (64 ciphertexts); M = 32: wave w takes row tile w%2 and
// column tiles 2(w/2), 2(w/2)+1 (128 ciphertexts).
// ---------------------------------------------------------------------------
template <int N>
struct GemmShape {
    static constexpr int M = Ring<N>::M, R = Ring<N>::R, NB = 2 * R;
    static constexpr int RT = 2 * M / 32;
    static constexpr int CTW = 2;                           // column tiles per wave
    static constexpr int CT_WG = 32 * CTW * (4 / RT);       // ciphertexts per workgroup
    static constexpr int KEY_ROW = 2 * M;                   // bytes per limb row
    static constexpr int CHUNK_BYTES = SPLIT_PC * 2 * SPLIT_LIMBS * KEY_ROW;
};

template <int N>
__global__ __launch_bounds__(SPLIT_THREADS, 2) void split_gemm_kernel(
    const int8_t *__restrict__ key_g, const int8_t *__restrict__ dtiles,
    int64_t *__restrict__ bins_out, int count, int P_pad)
{
    using G = GemmShape<N>;
    constexpr int M = G::M, NB = G::NB, KR = G::KEY_ROW;
    __shared__ __align__(16) int8_t kl[G::CHUNK_BYTES + 16];
    const int bin = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rt = w % G::RT;
    const int ct_tile0 = blockIdx.y * (G::CT_WG / 32) + (w / G::RT) * G::CTW;
    const int row = rt * 32 + (lane & 31);           // (a, k) row of this lane's A data
    const int a = row / M, kout = row % M;
    const int khalf = lane >> 5;
    const int KC = P_pad * (M / 16);
    const int8_t *keyb = key_g + (size_t)bin * P_pad * 2 * SPLIT_LIMBS * KR;

    v16i_s acc[SPLIT_LIMBS][G::CTW];
#pragma unroll
    for (int u = 0; u < SPLIT_LIMBS; u++)
#pragma unroll
        for (int c = 0; c < G::CTW; c++)
            acc[u][c] = (v16i_s){0};

    // digit tiles of column tile c, limb s: [kc / 8][ciphertext][kc % 8][16 bytes]
    const int8_t *bt[G::CTW][2];
#pragma unroll
    for (int c = 0; c < G::CTW; c++)
#pragma unroll
        for (int s = 0; s < 2; s++)
            bt[c][s] = dtiles + (((size_t)(ct_tile0 + c) * NB + bin) * 2 + s) * KC * 512 +
                       (size_t)(lane & 31) * 128;
    const int KS = P_pad * (M / 32);                 // k steps of 32
    auto load_b = [&](int ks, v4i_s (&Bx)[G::CTW][2]) {
        const int kc = 2 * ks + khalf;
        const size_t off = (size_t)(kc >> 3) * 32 * 128 + (kc & 7) * 16;
#pragma unroll
        for (int c = 0; c < G::CTW; c++)
#pragma unroll
            for (int s = 0; s < 2; s++)
                Bx[c][s] = *reinterpret_cast<const v4i_s *>(bt[c][s] + off);
    };
    // one k step: Toeplitz rows of the key bin from LDS (five limbs), 18 MFMAs
    auto kstep = [&](int kl_, const v4i_s (&B)[G::CTW][2]) {
        const int pl = kl_ / (M / 32), h = kl_ % (M / 32);
        const int x0 = M - 1 - kout + h * 32 + khalf * 16;
        const int8_t *rowp = kl + (pl * 2 + a) * SPLIT_LIMBS * KR;
        v4i_s A[SPLIT_LIMBS];
#pragma unroll
        for (int t = 0; t < SPLIT_LIMBS; t++) {
            const uint32_t *wp = reinterpret_cast<const uint32_t *>(rowp + t * KR + (x0 & ~3));
            const uint32_t w0 = wp[0], w1 = wp[1], w2 = wp[2], w3 = wp[3], w4 = wp[4];
            const uint32_t sh = (uint32_t)(x0 & 3);
            A[t][0] = (int)__builtin_amdgcn_alignbyte(w1, w0, sh);
            A[t][1] = (int)__builtin_amdgcn_alignbyte(w2, w1, sh);
            A[t][2] = (int)__builtin_amdgcn_alignbyte(w3, w2, sh);
            A[t][3] = (int)__builtin_amdgcn_alignbyte(w4, w3, sh);
        }
#pragma unroll
        for (int c = 0; c < G::CTW; c++) {
#pragma unroll
            for (int t = 0; t < SPLIT_LIMBS; t++)
                acc[t][c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[t], B[c][0], acc[t][c], 0, 0, 0);
#pragma unroll
            for (int t = 0; t + 1 < SPLIT_LIMBS; t++)
                acc[t + 1][c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[t], B[c][1], acc[t + 1][c], 0, 0, 0);
        }
    };

    // B fragments run one k step ahead of the MFMAs (two register sets)
    v4i_s B0[G::CTW][2], B1[G::CTW][2];
    load_b(0, B0);
    constexpr int KL = SPLIT_PC * (M / 32);          // k steps per key chunk (even)
    for (int pc0 = 0; pc0 < P_pad; pc0 += SPLIT_PC) {
        __syncthreads();
        {
            const v4i_s *src = reinterpret_cast<const v4i_s *>(keyb + (size_t)pc0 * 2 * SPLIT_LIMBS * KR);
            v4i_s *dst = reinterpret_cast<v4i_s *>(kl);
            for (int e = threadIdx.x; e < G::CHUNK_BYTES / 16; e += SPLIT_THREADS)
                dst[e] = src[e];
        }
        __syncthreads();
        const int ks0 = pc0 * (M / 32);
#pragma unroll 1
        for (int k2 = 0; k2 < KL; k2 += 2) {
            load_b(ks0 + k2 + 1, B1);                  // ks0 + k2 + 1 < KS: KL is even
            kstep(k2, B0);
            if (ks0 + k2 + 2 < KS)
                load_b(ks0 + k2 + 2, B0);
            kstep(k2 + 1, B1);
        }
    }
    // C[(a,k), ct] = sum_u acc_u 2^(8u); register r of lane: row 8(r/4) + 4(lane/32) + r%4
#pragma unroll
    for (int c = 0; c < G::CTW; c++) {
        const int ctn = (ct_tile0 + c) * 32 + (lane & 31);
        if (ctn >= count)
            continue;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int rr = rt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int aa = rr / M, kk = rr % M;
            uint64_t val = 0;
#pragma unroll
            for (int u = 0; u < SPLIT_LIMBS; u++)
                val += (uint64_t)(int64_t)acc[u][c][r] << (8 * u);
            bins_out[(((size_t)ctn * 2 + aa) * NB + bin) * M + kk] = (int64_t)val;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static bool split_degree_ok(int64_t degree)
{
    return degree == 512 || degree == 1024 || degree == 2048;
}

static void ring_of(int64_t degree, int *M, int *R)
{
    switch (degree) {
    case 512: *M = 32; *R = 16; break;
    case 1024: *M = 32; *R = 32; break;
    default: *M = 64; *R = 32; break;
    }
}

static int64_t pbc_pad(int64_t l, int64_t m) { return (2 * l * m + SPLIT_PC - 1) / SPLIT_PC * SPLIT_PC; }

static int64_t ct_per_wg(int64_t degree) { return degree <= 1024 ? 128 : 64; }

// bytes of the step scratch per ciphertext: digit tiles + product bins
static int64_t per_ct_bytes(int64_t degree, int64_t l, int64_t m)
{
    int M, R;
    ring_of(degree, &M, &R);
    const int64_t NB = 2 * R;
    return NB * 2 * pbc_pad(l, m) * M + 2 * NB * M * 8;
}

template <int N>
static int launch_key(hipStream_t s, const double *bsk, int n_iter, int L, int m, int P_pad,
                      int8_t *key)
{
    const int64_t polys = (int64_t)n_iter * 2 * L * 2 * m;
    hipLaunchKernelGGL(split_key_kernel<N>, dim3((unsigned)polys), dim3(SPLIT_THREADS), 0, s,
                       reinterpret_cast<const cplx *>(bsk), L, m, P_pad, key);
    if (int st = launch_status("split key"))
        return st;
    const int P = 2 * L * m;
    if (P_pad > P) {
        constexpr int M = Ring<N>::M, R = Ring<N>::R;
        const int64_t blocks = (int64_t)n_iter * 2 * R;
        hipLaunchKernelGGL(split_key_pad_kernel, dim3(1024), dim3(256), 0, s, key, blocks, P, P_pad,
                           2 * SPLIT_LIMBS * 2 * M);
        return launch_status("split key padding");
    }
    return BSLSTM_OK;
}

template <int N, int L>
static int run_chunks(hipStream_t s, const int64_t *ct_init, int64_t ct_stride, const int8_t *key,
                      const int32_t *indices, int64_t *output, int64_t batch, SplitShape sh,
                      int8_t *ws, int64_t chunk)
{
    using G = GemmShape<N>;
    constexpr int M = G::M, NB = G::NB;
    const int64_t dt_bytes = (chunk / 32) * (int64_t)NB * 2 * sh.P_pad * (M / 16) * 512;
    int8_t *dtiles = ws;
    int64_t *bins = reinterpret_cast<int64_t *>(ws + dt_bytes);
    const int64_t key_step = (int64_t)NB * sh.P_pad * 2 * SPLIT_LIMBS * G::KEY_ROW;
    for (int64_t c0 = 0; c0 < batch; c0 += chunk) {
        const int cnt = (int)std::min<int64_t>(chunk, batch - c0);
        const int64_t *cti = ct_init + (ct_stride ? c0 * ct_stride : 0);
        const int32_t *idx = indices + c0 * sh.n_iter * sh.m;
        int64_t *outc = output + c0 * 2 * N;
        const dim3 ggrid((unsigned)NB, (unsigned)((cnt + G::CT_WG - 1) / G::CT_WG));
        for (int g = 0; g <= sh.n_iter; g++) {
            hipLaunchKernelGGL((split_step_kernel<N, L>), dim3((unsigned)cnt), dim3(SPLIT_THREADS), 0,
                               s, cti, ct_stride, bins, idx, dtiles, outc, cnt, g, sh);
            if (g == sh.n_iter)
                break;
            hipLaunchKernelGGL(split_gemm_kernel<N>, ggrid, dim3(SPLIT_THREADS), 0, s,
                               key + (size_t)g * key_step, dtiles, bins, cnt, sh.P_pad);
        }
        if (int st = launch_status("split blind rotation"))
            return st;
    }
    return BSLSTM_OK;
}

template <int N>
static int dispatch_split_l(int64_t l, hipStream_t s, const int64_t *ct, int64_t st,
                            const int8_t *key, const int32_t *idx, int64_t *out, int64_t batch,
                            SplitShape sh, int8_t *ws, int64_t chunk)
{
    switch (l) {
    case 1: return run_chunks<N, 1>(s, ct, st, key, idx, out, batch, sh, ws, chunk);
    case 2: return run_chunks<N, 2>(s, ct, st, key, idx, out, batch, sh, ws, chunk);
    case 3: return run_chunks<N, 3>(s, ct, st, key, idx, out, batch, sh, ws, chunk);
    case 4: return run_chunks<N, 4>(s, ct, st, key, idx, out, batch, sh, ws, chunk);
    default: return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported l");
    }
}

static int split_dims(const int64_t *bsk_dims, int64_t l, int64_t degree, int64_t *n_iter,
                      int64_t *m)
{
    if (!split_degree_ok(degree))
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "Unsupported degree for the split form (512, 1024 or 2048)");
    if (l < 1 || l > 4)
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported l");
    if (!bsk_dims)
        return fail(BSLSTM_INVALID_ARGUMENT, "bsk_dims is NULL");
    if (bsk_dims[1] != 2 || bsk_dims[2] != l || bsk_dims[3] != 2 || bsk_dims[5] != degree / 2)
        return fail(BSLSTM_INVALID_ARGUMENT, "bsk trailing dims do not match (n_iter,2,l,2,m,N/2)");
    if (bsk_dims[0] < 1 || bsk_dims[4] < 1 || bsk_dims[4] > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "bsk n_iter, m out of range");
    *n_iter = bsk_dims[0];
    *m = bsk_dims[4];
    return BSLSTM_OK;
}

}  // namespace bslstm

using namespace bslstm;

extern "C" int64_t bslstm_mul_bootstrap_split_key_bytes(const int64_t bsk_dims[6], int64_t degree)
{
    int64_t n_iter, m;
    if (!bsk_dims || split_dims(bsk_dims, bsk_dims[2], degree, &n_iter, &m) != BSLSTM_OK)
        return -1;
    int M, R;
    ring_of(degree, &M, &R);
    return n_iter * 2 * R * pbc_pad(bsk_dims[2], m) * 2 * SPLIT_LIMBS * 2 * M;
}

extern "C" int bslstm_mul_bootstrap_split_prepare_key(void *stream, const double *bsk,
                                                      const int64_t bsk_dims[6], int64_t l,
                                                      int64_t degree, int8_t *key)
{
    clear_error();
    int64_t n_iter, m;
    if (int st = split_dims(bsk_dims, l, degree, &n_iter, &m))
        return st;
    if (!bsk || !key)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    if (n_iter * 2 * l * 2 * m > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "key too large");
    hipStream_t s = (hipStream_t)stream;
    const int P_pad = (int)pbc_pad(l, m);
    switch (degree) {
    case 512: return launch_key<512>(s, bsk, (int)n_iter, (int)l, (int)m, P_pad, key);
    case 1024: return launch_key<1024>(s, bsk, (int)n_iter, (int)l, (int)m, P_pad, key);
    default: return launch_key<2048>(s, bsk, (int)n_iter, (int)l, (int)m, P_pad, key);
    }
}

extern "C" int64_t bslstm_mul_bootstrap_split_workspace_bytes(int64_t batch, const int64_t bsk_dims[6],
                                                              int64_t degree, int64_t chunk)
{
    int64_t n_iter, m;
    if (batch < 0 || !bsk_dims || split_dims(bsk_dims, bsk_dims[2], degree, &n_iter, &m) != BSLSTM_OK)
        return -1;
    const int64_t wg = ct_per_wg(degree);
    if (chunk <= 0)
        chunk = 512;
    chunk = std::min(chunk, std::max<int64_t>(batch, 1));
    chunk = (chunk + wg - 1) / wg * wg;
    return chunk * per_ct_bytes(degree, bsk_dims[2], m);
}

extern "C" int bslstm_mul_bootstrap_split(void *stream, const int64_t *ct_init, int64_t ct_elems,
                                          const int8_t *split_key, const int64_t bsk_dims[6],
                                          const int32_t *indices, int64_t idx_elems,
                                          int64_t *output, int64_t batch, int64_t log2_beta,
                                          int64_t log2_q, int64_t l, int64_t degree,
                                          void *workspace, int64_t workspace_bytes)
{
    clear_error();
    int64_t n_iter, m;
    if (int st = split_dims(bsk_dims, l, degree, &n_iter, &m))
        return st;
    if (log2_q < 2 || log2_q > 32)
        return fail(BSLSTM_INVALID_ARGUMENT, "the split form needs 2 <= log2_q <= 32");
    if (log2_beta < 1 || log2_beta > 7 || l * log2_beta > log2_q)
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "the split form needs 1 <= log2_beta <= 7 (int8 digits), l*log2_beta <= log2_q");
    if (batch < 0 || batch > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "batch out of range");
    int64_t stride;
    if (ct_elems == 2 * degree)
        stride = 0;
    else if (ct_elems == batch * 2 * degree)
        stride = 2 * degree;
    else
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "ct_init has %lld elements; expected 2N=%lld or batch*2N=%lld",
                    (long long)ct_elems, (long long)(2 * degree), (long long)(batch * 2 * degree));
    if (idx_elems != batch * n_iter * m)
        return fail(BSLSTM_INVALID_ARGUMENT, "indices has %lld elements; expected %lld",
                    (long long)idx_elems, (long long)(batch * n_iter * m));
    if (batch == 0)
        return BSLSTM_OK;
    if (!ct_init || !split_key || !indices || !output || !workspace)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const int64_t wg = ct_per_wg(degree);
    const int64_t per = per_ct_bytes(degree, l, m);
    int64_t chunk = workspace_bytes / per / wg * wg;
    if (chunk < wg)
        return fail(BSLSTM_INVALID_ARGUMENT, "workspace of %lld bytes; need at least %lld",
                    (long long)workspace_bytes, (long long)(wg * per));
    chunk = std::min(chunk, (batch + wg - 1) / wg * wg);
    const SplitShape sh{(int)n_iter, (int)m, (int)(2 * l * m), (int)pbc_pad(l, m), (int)log2_beta,
                        (int)log2_q};
    hipStream_t s = (hipStream_t)stream;
    int8_t *ws = reinterpret_cast<int8_t *>(workspace);
    switch (degree) {
    case 512: return dispatch_split_l<512>(l, s, ct_init, stride, split_key, indices, output, batch, sh, ws, chunk);
    case 1024: return dispatch_split_l<1024>(l, s, ct_init, stride, split_key, indices, output, batch, sh, ws, chunk);
    default: return dispatch_split_l<2048>(l, s, ct_init, stride, split_key, indices, output, batch, sh, ws, chunk);
    }
}
