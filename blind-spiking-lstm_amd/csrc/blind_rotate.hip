// Collapsed (key-unrolled) blind rotation on gfx950, FP64 transform forms.
//
// Semantics are those of the reference kernel (csrc/mul_bootstrap.cuh:20-136):
// per step g
//   dec[bc]  = signed gadget digits of the accumulator (csrc/dec_fft.cuh:4-31),
//              bc = 2*level + component
//   prod[a]  = sum_bc (sum_p BSK[g,a,bc,p] * FFT(X^idx[g,p])) * FFT(dec[bc])
//   ct[a]    = round_half_even(centred_mod(IFFT(prod[a]), q))
//
// Design (MI355X-first, not a translation):
//  * Folded transform: each real length-N polynomial is one N/2-point complex
//    FFT (u_k = (p_k - i p_{k+N/2}) zeta^k), half the work of the N-point
//    complex transform on real data.
//  * T = N/8 threads per ciphertext; each owns the coefficients k = t + i*T
//    (i < 8) of both accumulator rows, in registers, for the whole loop.  The
//    first forward radix pass reads them directly and the last inverse pass
//    writes them back, so the accumulator never visits LDS.
//  * Stockham radix-4 passes (plus one radix-2 pass when log2(N/2) is odd)
//    exchange through LDS, one gadget level (the A- and B-digit polynomials)
//    at a time.
//  * The multiply-accumulate runs per level and per bin pair (r, N/2-1-r):
//    these U-bins are the reference's FFT bins 2r and 2r+1, so every BSK read
//    is 32 contiguous bytes per lane; the next pattern's BSK values are
//    prefetched while the current one is consumed, keeping eight 16-byte
//    loads in flight per lane.
//  * CT ciphertexts per workgroup.  CT = 1 is the plain single-launch form
//    (one workgroup per ciphertext, as the reference).  CT = 2 is the batched
//    form: a thread multiplies each BSK value it loads into both
//    ciphertexts, halving the key traffic from L2.  RG = 2 (with CT = 1) is
//    the latency form: the ciphertext's two accumulator rows sit in two
//    thread groups, each transforming its own row's digit polynomials, and
//    the multiply-accumulate spreads over both groups, so a step runs on
//    twice the threads.  All forms run identical arithmetic per ciphertext,
//    so their outputs are bit-identical.
//  * Roots of unity come from a half-period table in LDS (N complex),
//    filled once per workgroup with the device sincospi: no 2N x N/2
//    rotation table is read.
#include <type_traits>

#include "br_common.h"
#include "capi_internal.h"

namespace bslstm {

template <int N, int L, int CT, int RG, int ACC, int MP>
__global__ __launch_bounds__(N / 8 * CT * RG, 2) void br_kernel(const int64_t *__restrict__ ct_init,
                                                        int64_t ct_stride,
                                                        const cplx *__restrict__ bsk,
                                                        const int32_t *__restrict__ indices,
                                                        int64_t *__restrict__ out, int64_t batch,
                                                        int n_iter, int m_arg, int log2_beta,
                                                        int log2_q, KeySlots ks)
{
    // MP > 0: the pattern count is a compile-time constant (unrolled loops)
    const int m = MP > 0 ? MP : m_arg;
    constexpr int M = Fold<N>::M, T = Fold<N>::T, BC = 2 * L;
    constexpr int NR = 2 / RG;                   // accumulator rows per thread group
    constexpr int PPT = 2 / (CT * RG);           // bin pairs per thread in the MAC
    constexpr int NA = N <= 2048 ? 2 : 1;        // output rows sharing one root lookup
    static_assert((CT == 1 || CT == 2) && (RG == 1 || RG == 2) && CT * RG <= 2,
                  "one or two ciphertexts per workgroup, or one split over two row groups");
    static_assert(RG == 1 || Wave<N>::OK, "row groups need the wave-local transform");
    // 512-thread workgroups (two ciphertexts, or two row groups): full-period
    // root table (no sign flips; at most 128 KiB of LDS in all at N = 2048)
    constexpr bool FULL = CT * RG == 2 && N <= 2048;
    // levels transformed per phase: the row-group form transforms two gadget
    // levels of its row at once (two independent transforms per thread, and
    // half the phases per step); the others one level of both rows
    constexpr int LP = RG == 2 ? 2 : 1;
    __shared__ cplx buf[CT * 2 * M * LP];
    __shared__ cplx ztab[RootTab<N, FULL>::SIZE];

    const int grp = threadIdx.x / T;             // thread group (transforms)
    const int t = threadIdx.x - grp * T;         // thread index within the group
    const int cs = grp / RG;                     // ciphertext slot of the group
    const int row0 = (grp - cs * RG) * NR;       // first accumulator row the group owns
    const int u = threadIdx.x;                   // MAC unit
    const int64_t first = (int64_t)blockIdx.x * CT;
    int64_t ct_of[CT];
#pragma unroll
    for (int c = 0; c < CT; c++)
        ct_of[c] = first + c < batch ? first + c : batch - 1;   // tail: recompute the last, store nothing

    if (ks.slots) {
        int sl = ks.slots[ct_of[0]];
        sl = sl < 0 ? 0 : (sl >= ks.count ? ks.count - 1 : sl);
        bsk += (size_t)sl * ks.stride;
    }
    fill_root_table<N, FULL>(ztab);

    // accumulator rows in registers: x[c][i] = ct[(row0 + c)*N + t + i*T]
    constexpr bool T32 = ACC != ACC_64;
    using acc_t = typename std::conditional<T32, uint32_t, uint64_t>::type;
    const int shift = ACC == ACC_HI ? log2_q - 32 : 0;   // register = class >> shift
    acc_t x[NR][8];
    const int64_t *src = ct_init + ct_of[cs] * ct_stride + (size_t)row0 * N;
#pragma unroll
    for (int c = 0; c < NR; c++)
#pragma unroll
        for (int i = 0; i < 8; i++)
            x[c][i] = (acc_t)((uint64_t)src[c * N + t + i * T] >> shift);
    __syncthreads();

    // decomposition offset (dec_fft.cuh:7-17)
    uint64_t off = 0;
    for (int i = 1; i <= L; i++)
        off += 1ULL << (log2_q - i * log2_beta + log2_beta - 1);
    {
        const int rs = log2_q - L * log2_beta - 1;
        if (rs >= 0)
            off += 1ULL << rs;
    }
    const acc_t offs = (acc_t)(off >> shift);      // off < 2^log2_q, a multiple of 2^shift
    const acc_t dmask = (acc_t)((1ULL << log2_beta) - 1);
    const int half_beta = 1 << (log2_beta - 1);

    const double q = ldexp(1.0, log2_q);
    const double inv_q = ldexp(1.0, -log2_q);
    const double inv_m = 1.0 / (double)M;
    const size_t step_stride = (size_t)2 * BC * m * M;

    // L2 warm-up of the next step's key slice (512-thread forms, one workgroup
    // per CU).  Workgroups dispatched to one XCD (blockIdx.x mod 8) run the
    // key in near lockstep; the P resident ones split the slice's 128-byte
    // lines, and each thread touches one word of up to NPF lines right after
    // its multiply-accumulate, so the lines are in L2 when the next step's MAC
    // reads them.  The words are folded into `warm`, which is consumed one
    // step later (the loads are long complete by then) and only stored under a
    // condition that never holds.  (In the 256-thread form the two workgroups
    // per CU drift apart and the extra lines evict each other's slices.)
    // NPF: the latency form runs few workgroups per XCD (P = 4 at 32
    // ciphertexts), each then covering up to 2048 lines of a 1 MiB slice.
    constexpr bool WARM = CT * RG == 2;
    constexpr int NPF = RG == 2 ? 4 : 2;
    const int P = min(32, max(1, (int)((gridDim.x + 7) / 8)));
    const int lines = (int)(step_stride / 8);          // 128-byte lines per key slice
    const int lpw = (lines + P - 1) / P;               // lines per workgroup
    const int line0 = ((int)(blockIdx.x >> 3) % P) * lpw;
    const int line_end = min(lines, line0 + lpw);
    uint32_t warm = 0, pf0 = 0, pf1 = 0, pf2 = 0, pf3 = 0;
    // forward target of this group: RG = 2: row0's two levels [level][M];
    // RG = 1: both rows of ciphertext cs [row][M].  The product spectra of
    // ciphertext c sit at buf[c*2M + a*M] for the inverse.
    cplx *fwd = buf + (RG == 2 ? row0 * 2 * M : cs * 2 * M);
    cplx *mybuf = buf + cs * 2 * M + row0 * M;

#pragma unroll 1
    for (int g = 0; g < n_iter; g++) {
        const cplx *bsk_g = bsk + (size_t)g * step_stride;
        const int32_t *idp[CT];
#pragma unroll
        for (int c = 0; c < CT; c++)
            idp[c] = indices + (ct_of[c] * n_iter + g) * m;

        cplx prod[2][PPT][CT][2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int pp = 0; pp < PPT; pp++)
#pragma unroll
                for (int c = 0; c < CT; c++)
                    prod[a][pp][c][0] = prod[a][pp][c][1] = {0.0, 0.0};

#pragma unroll 1
        for (int lev0 = 0; lev0 < L; lev0 += LP) {
            // ---- digits (dec_fft.cuh:19-29), fold, forward transform: polynomial
            // k is row k at level lev0 (RG = 1) or row0 at level lev0 + k (RG = 2)
            cplx v[2][4];
#pragma unroll
            for (int k = 0; k < 2; k++) {
                // (past the last level -- the odd-l case, transformed by no one --
                // the last level's shift keeps the arithmetic defined)
                const int lev = min(lev0 + (RG == 2 ? k : 0), L - 1);
                const int c = RG == 2 ? 0 : k;
                const int sh = log2_q - (lev + 1) * log2_beta - shift;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int lo = (int)(((acc_t)(x[c][r] + offs) >> sh) & dmask) - half_beta;
                    int hi = (int)(((acc_t)(x[c][r + 4] + offs) >> sh) & dmask) - half_beta;
                    v[k][r] = cmul({(double)lo, -(double)hi}, zeta_pow<N, FULL>(ztab, t + r * T));   // (lo - i*hi) * zeta^k
                }
            }
            if (RG == 2 && lev0 + 1 == L) {
                // odd l: the last phase transforms one level
                cplx v1[1][4] = {{v[0][0], v[0][1], v[0][2], v[0][3]}};
                if constexpr (Wave<N>::OK)
                    forward_wave<N, FULL, 1>(v1, fwd, ztab, t);
                else
                    forward_pair<N, FULL, 1>(v1, fwd, ztab, t);
            } else {
                if constexpr (Wave<N>::OK)
                    forward_wave<N, FULL, 2>(v, fwd, ztab, t);
                else
                    forward_pair<N, FULL, 2>(v, fwd, ztab, t);
            }

#pragma unroll 1
            for (int l2 = 0; l2 < LP; l2++) {
            const int lev = lev0 + l2;
            if (lev >= L)
                break;
            // ---- external product, level lev (mul_bootstrap.cuh:42-67)
#pragma unroll 1
            for (int pp = 0; pp < PPT; pp++) {
                const int r = CT * RG == 1 ? u + pp * T : u;   // r < M/2; partner M-1-r
#pragma unroll 1
                for (int a0 = 0; a0 < 2; a0 += NA) {
                    cplx W[NA][2][CT][2];
#pragma unroll
                    for (int a = 0; a < NA; a++)
#pragma unroll
                        for (int b2 = 0; b2 < 2; b2++)
#pragma unroll
                            for (int c = 0; c < CT; c++)
                                W[a][b2][c][0] = W[a][b2][c][1] = {0.0, 0.0};
                    // ping-pong prefetch: the loads of pattern p+1 fly while p is consumed
                    cplx bkA[NA][2][2], bkB[NA][2][2];
                    load_pair<L, NA>(bkA, bsk_g, 2 * r, a0, lev, 0, m, M);
#pragma unroll 1
                    for (int p = 0; p < m; p += 2) {
                        load_pair<L, NA>(bkB, bsk_g, 2 * r, a0, lev, p + 1 < m ? p + 1 : p, m, M);
                        mac_pattern<N, CT, NA, FULL>(W, bkA, idp, p, r, ztab);
                        if (p + 1 < m) {
                            load_pair<L, NA>(bkA, bsk_g, 2 * r, a0, lev, p + 2 < m ? p + 2 : p + 1, m, M);
                            mac_pattern<N, CT, NA, FULL>(W, bkB, idp, p + 1, r, ztab);
                        }
                    }
                    cplx acc[NA][CT][2];
#pragma unroll
                    for (int c = 0; c < CT; c++) {
#pragma unroll
                        for (int a = 0; a < NA; a++)
                            acc[a][c][0] = acc[a][c][1] = {0.0, 0.0};
#pragma unroll
                        for (int b2 = 0; b2 < 2; b2++) {
                            const int dofs = RG == 2 ? b2 * 2 * M + l2 * M : c * 2 * M + b2 * M;
                            const cplx d0 = buf[dofs + phys<N>(r)];
                            const cplx d1 = buf[dofs + phys<N>(M - 1 - r)];
#pragma unroll
                            for (int a = 0; a < NA; a++) {
                                acc[a][c][0] = cfma(W[a][b2][c][0], d0, acc[a][c][0]);
                                acc[a][c][1] = cfma(W[a][b2][c][1], d1, acc[a][c][1]);
                            }
                        }
                    }
                    // (a0, pp) are uniform loop counters: select the accumulator
                    // with compile-time indices so it stays in registers
#pragma unroll
                    for (int aa = 0; aa < 2; aa += NA)
#pragma unroll
                        for (int qq = 0; qq < PPT; qq++)
                            if (aa == a0 && qq == pp)
#pragma unroll
                                for (int a = 0; a < NA; a++)
#pragma unroll
                                    for (int c = 0; c < CT; c++) {
                                        prod[aa + a][qq][c][0] = cadd(prod[aa + a][qq][c][0], acc[a][c][0]);
                                        prod[aa + a][qq][c][1] = cadd(prod[aa + a][qq][c][1], acc[a][c][1]);
                                    }
                }
            }
            }
        }

        warm ^= pf0 ^ pf1;
        pf0 = pf1 = 0;
        if constexpr (NPF == 4) {
            warm ^= pf2 ^ pf3;
            pf2 = pf3 = 0;
        }
        if (WARM && g + 1 < n_iter) {
            const uint32_t *nxt = reinterpret_cast<const uint32_t *>(bsk + (size_t)(g + 1) * step_stride);
            const int j0 = line0 + (int)threadIdx.x, j1 = j0 + (int)blockDim.x;
            if (j0 < line_end)
                pf0 = nxt[(size_t)j0 * 32];
            if (j1 < line_end)
                pf1 = nxt[(size_t)j1 * 32];
            if constexpr (NPF == 4) {
                const int j2 = j1 + (int)blockDim.x, j3 = j2 + (int)blockDim.x;
                if (j2 < line_end)
                    pf2 = nxt[(size_t)j2 * 32];
                if (j3 < line_end)
                    pf3 = nxt[(size_t)j3 * 32];
            }
        }

        // ---- inverse transforms, untwist, centred mod q, round (mul_bootstrap.cuh:70-93)
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < PPT; pp++) {
            const int r = CT * RG == 1 ? u + pp * T : u;
#pragma unroll
            for (int c = 0; c < CT; c++)
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    buf[c * 2 * M + a * M + phys<N>(r)] = prod[a][pp][c][0];
                    buf[c * 2 * M + a * M + phys<N>(M - 1 - r)] = prod[a][pp][c][1];
                }
        }
        __syncthreads();
        cplx w[NR][4];
        if constexpr (Wave<N>::OK)
            inverse_wave<N, FULL, NR>(w, mybuf, ztab, t);
        else
            inverse_pair<N, FULL, NR>(w, mybuf, ztab, t);
#pragma unroll
        for (int a = 0; a < NR; a++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const cplx z = cmul(w[a][r], zeta_pow<N, FULL>(ztab, -(t + r * T)));
                double lo = z.re * inv_m;
                double hi = -z.im * inv_m;
                if constexpr (T32) {
                    // rint(val) - q*round_half_up(rint(val)/q) lies in [-q/2, q/2)
                    // and is the class of the reference's round(val - q*rint(val/q))
                    // (the subtraction is exact for |val| < 2^52)
                    // (also exact for |val| up to 2^80 at log2_q = 64: the products
                    // are then integer-valued and lo/q + 0.5 is exact)
                    lo = rint(lo);
                    hi = rint(hi);
                    lo = lo - q * floor(fma(lo, inv_q, 0.5));
                    hi = hi - q * floor(fma(hi, inv_q, 0.5));
                    if (ACC == ACC_HI && g == n_iter - 1 && first + cs < batch) {
                        int64_t *dst = out + (first + cs) * 2 * N + (size_t)(row0 + a) * N;
                        dst[t + r * T] = (int64_t)lo;
                        dst[t + (r + 4) * T] = (int64_t)hi;
                    }
                    x[a][r] = (acc_t)((uint64_t)(int64_t)lo >> shift);
                    x[a][r + 4] = (acc_t)((uint64_t)(int64_t)hi >> shift);
                } else {
                    lo = lo - q * rint(lo * inv_q);
                    hi = hi - q * rint(hi * inv_q);
                    lo = rint(lo);
                    hi = rint(hi);
                    if (log2_q >= 64) {   // +2^63 is the class of -2^63
                        lo = lo >= 0x1p63 ? -0x1p63 : lo;
                        hi = hi >= 0x1p63 ? -0x1p63 : hi;
                    }
                    x[a][r] = (acc_t)(int64_t)lo;
                    x[a][r + 4] = (acc_t)(int64_t)hi;
                }
            }
        }
    }

    if constexpr (NPF == 4)
        warm ^= pf2 ^ pf3;
    if ((warm ^ pf0 ^ pf1) == 0x9E3779B9u && ks.count < 0)   // ks.count >= 1 always
        out[0] = (int64_t)warm;
    if (ACC != ACC_HI && first + cs < batch) {
        int64_t *dst = out + (first + cs) * 2 * N + (size_t)row0 * N;
        // T32: centred representative in [-2^(log2_q-1), 2^(log2_q-1))
        const int sh_out = 64 - log2_q;
#pragma unroll
        for (int c = 0; c < NR; c++)
#pragma unroll
            for (int i = 0; i < 8; i++)
                dst[c * N + t + i * T] =
                    T32 ? ((int64_t)((uint64_t)x[c][i] << sh_out) >> sh_out) : (int64_t)x[c][i];
    }
}

template <int N, int L, int CT, int RG, int MP>
static void launch_acc(hipStream_t stream, const int64_t *ct, int64_t ct_stride, const cplx *bk,
                       const int32_t *idx, int64_t *out, int64_t batch, int n_iter, int m,
                       int log2_beta, int log2_q, KeySlots ks)
{
    const dim3 grid((unsigned)((batch + CT - 1) / CT)), block(N / 8 * CT * RG);
    if (log2_q <= 32)
        hipLaunchKernelGGL((br_kernel<N, L, CT, RG, ACC_32, MP>), grid, block, 0, stream, ct,
                           ct_stride, bk, idx, out, batch, n_iter, m, log2_beta, log2_q, ks);
    else if (L * log2_beta + 1 <= 32)
        hipLaunchKernelGGL((br_kernel<N, L, CT, RG, ACC_HI, MP>), grid, block, 0, stream, ct,
                           ct_stride, bk, idx, out, batch, n_iter, m, log2_beta, log2_q, ks);
    else
        hipLaunchKernelGGL((br_kernel<N, L, CT, RG, ACC_64, MP>), grid, block, 0, stream, ct,
                           ct_stride, bk, idx, out, batch, n_iter, m, log2_beta, log2_q, ks);
}

template <int N, int L, int CT, int RG>
static int launch_br(hipStream_t stream, const int64_t *ct, int64_t ct_stride, const double *bsk,
                     const int32_t *idx, int64_t *out, int64_t batch, int n_iter, int m,
                     int log2_beta, int log2_q, KeySlots ks)
{
    const cplx *bk = reinterpret_cast<const cplx *>(bsk);
    // collapse 3 (8 patterns, the reference's default) at N = 1024, 2048 runs
    // with the pattern loop unrolled
    if ((N == 1024 || N == 2048) && m == 8)
        launch_acc<N, L, CT, RG, 8>(stream, ct, ct_stride, bk, idx, out, batch, n_iter, m, log2_beta,
                                    log2_q, ks);
    else
        launch_acc<N, L, CT, RG, 0>(stream, ct, ct_stride, bk, idx, out, batch, n_iter, m, log2_beta,
                                    log2_q, ks);
    return launch_status("blind rotation");
}

template <int N, int CT, int RG = 1>
static int dispatch_l(int64_t l, hipStream_t s, const int64_t *ct, int64_t st, const double *bsk,
                      const int32_t *idx, int64_t *out, int64_t batch, int n_iter, int m, int lb,
                      int lq, KeySlots ks)
{
    switch (l) {
    case 1: return launch_br<N, 1, CT, RG>(s, ct, st, bsk, idx, out, batch, n_iter, m, lb, lq, ks);
    case 2: return launch_br<N, 2, CT, RG>(s, ct, st, bsk, idx, out, batch, n_iter, m, lb, lq, ks);
    case 3: return launch_br<N, 3, CT, RG>(s, ct, st, bsk, idx, out, batch, n_iter, m, lb, lq, ks);
    case 4: return launch_br<N, 4, CT, RG>(s, ct, st, bsk, idx, out, batch, n_iter, m, lb, lq, ks);
    default: return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported l");
    }
}

// Host-side validation shared by all forms (mul_bootstrap.cuh:138-228 plus
// the shape checks the reference does not make).
int validate_blind_rotation(int64_t ct_elems, const double *bsk, const int64_t *bsk_dims,
                    const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                    const int32_t *indices, int64_t idx_elems, int64_t *output, int64_t batch,
                    int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree,
                    const int64_t *ct_init, int64_t *stride)
{
    switch (degree) {
    case 128: case 256: case 512: case 1024: case 2048: case 4096: break;
    default: return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported degree");
    }
    if (l < 1 || l > 4)
        return fail(BSLSTM_INVALID_ARGUMENT, "Unsupported l");
    if (!bsk_dims)
        return fail(BSLSTM_INVALID_ARGUMENT, "bsk_dims is NULL");
    const int64_t n_iter = bsk_dims[0], a_size = bsk_dims[1], b_size = bsk_dims[2],
                  c_size = bsk_dims[3], m = bsk_dims[4], d_size = bsk_dims[5];
    if (a_size != 2 || c_size != 2 || b_size != l || d_size != degree / 2)
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "bsk trailing dims (%lld,%lld,%lld,%lld,%lld,%lld) do not match "
                    "(n_iter,2,l=%lld,2,m,N/2=%lld)",
                    (long long)n_iter, (long long)a_size, (long long)b_size, (long long)c_size,
                    (long long)m, (long long)d_size, (long long)l, (long long)(degree / 2));
    if (n_iter < 1 || m < 1 || m > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "bsk n_iter=%lld, m=%lld out of range",
                    (long long)n_iter, (long long)m);
    if (log2_beta < 1 || log2_q < 2 || log2_q > 64 || l * log2_beta > log2_q)
        return fail(BSLSTM_INVALID_ARGUMENT, "need 1 <= log2_beta, l*log2_beta <= log2_q <= 64");
    if (batch < 0 || batch > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "batch out of range");
    if (ct_elems == 2 * degree)
        *stride = 0;
    else if (ct_elems == batch * 2 * degree)
        *stride = 2 * degree;
    else
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "ct_init has %lld elements; expected 2N=%lld or batch*2N=%lld",
                    (long long)ct_elems, (long long)(2 * degree), (long long)(batch * 2 * degree));
    if (idx_elems != batch * n_iter * m)
        return fail(BSLSTM_INVALID_ARGUMENT, "indices has %lld elements; expected %lld",
                    (long long)idx_elems, (long long)(batch * n_iter * m));
    if (all_rot_fft && (rot_rows != 2 * degree || rot_cols != degree / 2))
        return fail(BSLSTM_INVALID_ARGUMENT, "all_rot_fft must be (2N, N/2)");
    if (batch > 0 && (!ct_init || !bsk || !indices || !output))
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    return BSLSTM_OK;
}

// CT ciphertexts per workgroup; RG = 2 splits one ciphertext's two
// accumulator rows over two thread groups (N in 512..2048; other degrees run
// the one-group form, whose arithmetic is the same).
template <int CT, int RG = 1>
static int run(void *stream, const int64_t *ct_init, int64_t ct_elems, const double *bsk,
               const int64_t *bsk_dims, const double *all_rot_fft, int64_t rot_rows,
               int64_t rot_cols, const int32_t *indices, int64_t idx_elems, int64_t *output,
               int64_t batch, int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree,
               KeySlots ks = {nullptr, 0, 1})
{
    clear_error();
    int64_t st = 0;
    int rc = validate_blind_rotation(ct_elems, bsk, bsk_dims, all_rot_fft, rot_rows, rot_cols, indices, idx_elems,
                      output, batch, log2_beta, log2_q, l, degree, ct_init, &st);
    if (rc != BSLSTM_OK || batch == 0)
        return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int ni = (int)bsk_dims[0], mm = (int)bsk_dims[4], lb = (int)log2_beta, lq = (int)log2_q;
    switch (degree) {
    case 128: return dispatch_l<128, CT>(l, s, ct_init, st, bsk, indices, output, batch, ni, mm, lb, lq, ks);
    case 256: return dispatch_l<256, CT>(l, s, ct_init, st, bsk, indices, output, batch, ni, mm, lb, lq, ks);
    case 512: return dispatch_l<512, CT, RG>(l, s, ct_init, st, bsk, indices, output, batch, ni, mm, lb, lq, ks);
    case 1024: return dispatch_l<1024, CT, RG>(l, s, ct_init, st, bsk, indices, output, batch, ni, mm, lb, lq, ks);
    case 2048: return dispatch_l<2048, CT, RG>(l, s, ct_init, st, bsk, indices, output, batch, ni, mm, lb, lq, ks);
    default:
        // 4096 x 2 ciphertexts would need all 160 KiB of LDS: one per workgroup
        return dispatch_l<4096, 1>(l, s, ct_init, st, bsk, indices, output, batch, ni, mm, lb, lq, ks);
    }
}

}  // namespace bslstm

extern "C" int bslstm_mul_bootstrap(void *stream, const int64_t *ct_init, int64_t ct_elems,
                                    const double *bsk, const int64_t bsk_dims[6],
                                    const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                                    const int32_t *indices, int64_t idx_elems, int64_t *output,
                                    int64_t batch, int64_t log2_beta, int64_t log2_q, int64_t l,
                                    int64_t degree)
{
    return bslstm::run<1>(stream, ct_init, ct_elems, bsk, bsk_dims, all_rot_fft, rot_rows, rot_cols,
                          indices, idx_elems, output, batch, log2_beta, log2_q, l, degree);
}

extern "C" int bslstm_mul_bootstrap_batched(void *stream, const int64_t *ct_init, int64_t ct_elems,
                                            const double *bsk, const int64_t bsk_dims[6],
                                            const double *all_rot_fft, int64_t rot_rows,
                                            int64_t rot_cols, const int32_t *indices,
                                            int64_t idx_elems, int64_t *output, int64_t batch,
                                            int64_t log2_beta, int64_t log2_q, int64_t l,
                                            int64_t degree)
{
    return bslstm::run<2>(stream, ct_init, ct_elems, bsk, bsk_dims, all_rot_fft, rot_rows, rot_cols,
                          indices, idx_elems, output, batch, log2_beta, log2_q, l, degree);
}

extern "C" int bslstm_mul_bootstrap_slots(void *stream, const int64_t *ct_init, int64_t ct_elems,
                                          const double *bsk_pool, const int64_t bsk_dims[6],
                                          const int32_t *slots, int64_t slot_count,
                                          const int32_t *indices, int64_t idx_elems,
                                          int64_t *output, int64_t batch, int64_t log2_beta,
                                          int64_t log2_q, int64_t l, int64_t degree, int64_t form)
{
    if (!slots || slot_count < 1 || slot_count > 0x7fffffffLL || !bsk_dims)
        return bslstm::fail(BSLSTM_INVALID_ARGUMENT, "slots and 1 <= slot_count required");
    int64_t stride = 1;
    for (int i = 0; i < 6; i++)
        stride *= bsk_dims[i];
    const bslstm::KeySlots ks{slots, stride, (int32_t)slot_count};
    switch (form) {
    case 1:
        return bslstm::run<1>(stream, ct_init, ct_elems, bsk_pool, bsk_dims, nullptr, 0, 0, indices,
                              idx_elems, output, batch, log2_beta, log2_q, l, degree, ks);
    case 2:
        return bslstm::run<2>(stream, ct_init, ct_elems, bsk_pool, bsk_dims, nullptr, 0, 0, indices,
                              idx_elems, output, batch, log2_beta, log2_q, l, degree, ks);
    case 3:
        return bslstm::run<1, 2>(stream, ct_init, ct_elems, bsk_pool, bsk_dims, nullptr, 0, 0,
                                 indices, idx_elems, output, batch, log2_beta, log2_q, l, degree, ks);
    default:
        return bslstm::fail(BSLSTM_INVALID_ARGUMENT, "form must be 1, 2 or 3");
    }
}

extern "C" int bslstm_mul_bootstrap_rows(void *stream, const int64_t *ct_init, int64_t ct_elems,
                                         const double *bsk, const int64_t bsk_dims[6],
                                         const double *all_rot_fft, int64_t rot_rows,
                                         int64_t rot_cols, const int32_t *indices,
                                         int64_t idx_elems, int64_t *output, int64_t batch,
                                         int64_t log2_beta, int64_t log2_q, int64_t l,
                                         int64_t degree)
{
    return bslstm::run<1, 2>(stream, ct_init, ct_elems, bsk, bsk_dims, all_rot_fft, rot_rows,
                             rot_cols, indices, idx_elems, output, batch, log2_beta, log2_q, l,
                             degree);
}
