// Latency form of the collapsed blind rotation over two workgroups per
// ciphertext (the "pair" form), N = 1024 and 2048.  Same semantics and bit-identical
// outputs as the one-workgroup forms (blind_rotate.hip; the reference kernel
// csrc/mul_bootstrap.cuh:20-136).
//
// At a few dozen ciphertexts a CMux step on one CU is bound by reading the
// step's key slice from L2 into that CU (~148 GB/s per CU).  Here workgroup a
// of a pair keeps accumulator row a and, per step:
//   1. transforms the gadget digits of its row (its two thread groups take
//      alternate levels), into LDS;
//   2. publishes those spectra to its partner through L2: thread u writes
//      the bins it multiplies (u and M-1-u), which are the bins the
//      partner's thread u needs;
//   3. streams the key rows of output row a (half the key slice) against the
//      rotation roots -- work that does not need any digit spectrum, so the
//      partner's spectra travel meanwhile;
//   4. waits for the partner's flag, loads the partner row's spectra at its
//      bins into registers and combines: for each level, sum over the two
//      rows of (key patterns x spectrum), in the one-workgroup forms' order;
//   5. inverts its row (both thread groups, each on its own copy, so both
//      keep the row in registers for the next step's digits).
// The arithmetic per output coefficient is the one-workgroup forms' (same
// digits, transforms, pattern and level order), so the results are identical.
//
// Pairing: blocks b and b+8 (round-robin dispatch puts both on XCD b mod 8,
// sharing that XCD's L2); ciphertext (b/16)*8 + b%8, row (b/8)%2.  At the
// start the two workgroups exchange their XCC_ID: on one XCD the spectra go
// through plain stores and L1-bypassing loads; otherwise through coherent
// (system-scope) stores and loads.  Every wait on the partner is bounded
// (spin_limit polls); a wait that runs out sets bit 0 of the status word and
// the workgroup skips every later wait, so the kernel always ends, after at
// most one run-out wait.  All workgroups must be resident
// together: the host admits at most one pair per two CUs.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "br_common.h"
#include "capi_internal.h"

namespace bslstm {

// degrees of the pair form: one 512-thread workgroup per CU at N = 2048
// (160 KiB of LDS at l = 3), two 256-thread ones per CU at N = 1024
static bool pair_degree(int64_t n) { return n == 1024 || n == 2048; }

// Workspace (device memory the caller passes): [0] status word, flags, then
// the exchange buffers (double-buffered by step parity).
struct PairWs {
    uint32_t *status;    // bit 0: a wait ran out; bit 1: a pair straddled two XCDs
    uint32_t *flags;     // [pair][member]: steps published
    uint32_t *hello;     // [pair][member]: XCC_ID + 1
    cplx *xbuf;          // [pair][parity][member][L][M/2 threads][2 bins]
};

static int64_t pair_count(int64_t batch) { return (batch + 7) / 8 * 8; }

static int64_t pair_ws_bytes(int64_t batch, int64_t l, int64_t degree)
{
    const int64_t P = pair_count(batch);
    const int64_t head = (1 + 4 * P) * 4;
    const int64_t head_al = (head + 255) / 256 * 256;
    return head_al + P * 2 * 2 * l * (degree / 2) * 16;
}

static PairWs pair_ws(void *ws, int64_t batch)
{
    const int64_t P = pair_count(batch);
    uint32_t *w = reinterpret_cast<uint32_t *>(ws);
    const int64_t head = (1 + 4 * P) * 4;
    const int64_t head_al = (head + 255) / 256 * 256;
    return {w, w + 1, w + 1 + 2 * P,
            reinterpret_cast<cplx *>(reinterpret_cast<char *>(ws) + head_al)};
}

__global__ void pair_clear_kernel(uint32_t *w, int64_t words)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        w[i] = 0;
}

__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20); }

// bounded wait until *f >= target (one thread); false if it ran out.  SCOPE:
// __HIP_MEMORY_SCOPE_AGENT when both workgroups share one XCD's L2, else
// __HIP_MEMORY_SCOPE_SYSTEM.
template <int SCOPE>
__device__ __forceinline__ bool wait_at_least(const uint32_t *f, uint32_t target, int spin_limit)
{
    for (int it = 0; it < spin_limit; it++) {
        if (__hip_atomic_load(f, __ATOMIC_RELAXED, SCOPE) >= target)
            return true;
        __builtin_amdgcn_s_sleep(2);
    }
    return __hip_atomic_load(f, __ATOMIC_RELAXED, SCOPE) >= target;
}

template <int N, int L, int ACC, int MP>
__global__ __launch_bounds__(N / 4, 2) void br_pair_kernel(
    const int64_t *__restrict__ ct_init, int64_t ct_stride, const cplx *__restrict__ bsk,
    const int32_t *__restrict__ indices, int64_t *__restrict__ out, int64_t batch, int n_iter,
    int m_arg, int log2_beta, int log2_q, KeySlots ks, PairWs ws, int spin_limit)
{
    const int m = MP > 0 ? MP : m_arg;
    constexpr int M = Fold<N>::M, T = Fold<N>::T;
    // both rows' L spectra next to the full-period root table: 160 KiB at
    // N = 2048, l = 3 (80 KiB at N = 1024)
    static_assert(L <= 3, "the pair form holds l <= 3");
    constexpr bool FULL = true;
    __shared__ cplx own[L * M];                   // this row's digit spectra, then a product copy
    __shared__ cplx oth[L * M];                   // the partner row's digit spectra, then a copy
    __shared__ cplx ztab[RootTab<N, FULL>::SIZE];

    const int b = blockIdx.x;
    const int pair = (b / 16) * 8 + (b % 8);
    const int a = (b / 8) & 1;                    // the row this workgroup keeps
    const int grp = threadIdx.x / T, t = threadIdx.x - grp * T;
    const int u = threadIdx.x;                    // MAC: bin pair u
    const int64_t ct = pair < batch ? pair : batch - 1;
    if (ks.slots) {
        int sl = ks.slots[ct];
        sl = sl < 0 ? 0 : (sl >= ks.count ? ks.count - 1 : sl);
        bsk += (size_t)sl * ks.stride;
    }
    fill_root_table<N, FULL>(ztab);

    // ---- handshake (system scope: the placement is not known yet): are both
    // workgroups on one XCD?  The answer reaches the workgroup through own[0],
    // which the first step's transforms overwrite later.
    bool broken = false;                           // thread 0: a wait ran out
    if (threadIdx.x == 0) {
        const uint32_t me = xcc_id() + 1;
        __hip_atomic_store(&ws.hello[pair * 2 + a], me, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        const bool ok =
            wait_at_least<__HIP_MEMORY_SCOPE_SYSTEM>(&ws.hello[pair * 2 + 1 - a], 1u, spin_limit);
        const uint32_t other = __hip_atomic_load(&ws.hello[pair * 2 + 1 - a], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_SYSTEM);
        own[0].re = (ok && other == me) ? 1.0 : 0.0;
        broken = !ok;
        if (!ok)
            atomicOr(ws.status, 1u);
        else if (other != me)
            atomicOr(ws.status, 2u);
    }

    constexpr bool T32 = ACC != ACC_64;
    using acc_t = typename std::conditional<T32, uint32_t, uint64_t>::type;
    const int shift = ACC == ACC_HI ? log2_q - 32 : 0;
    acc_t x[8];
    {
        const int64_t *src = ct_init + ct * ct_stride + (size_t)a * N;
#pragma unroll
        for (int i = 0; i < 8; i++)
            x[i] = (acc_t)((uint64_t)src[t + i * T] >> shift);
    }
    __syncthreads();
    const bool local = own[0].re != 0.0;
    __syncthreads();

    uint64_t off = 0;
    for (int i = 1; i <= L; i++)
        off += 1ULL << (log2_q - i * log2_beta + log2_beta - 1);
    {
        const int rs = log2_q - L * log2_beta - 1;
        if (rs >= 0)
            off += 1ULL << rs;
    }
    const acc_t offs = (acc_t)(off >> shift);
    const acc_t dmask = (acc_t)((1ULL << log2_beta) - 1);
    const int half_beta = 1 << (log2_beta - 1);
    const double q = ldexp(1.0, log2_q);
    const double inv_q = ldexp(1.0, -log2_q);
    const double inv_m = 1.0 / (double)M;
    const size_t step_stride = (size_t)2 * 2 * L * m * M;   // complex per step (both rows)
    const size_t row_stride = step_stride / 2;              // this row's half of a step

    // L2 warm-up of this row's half of the next step's key slice, split by
    // line among the workgroups of one XCD that keep the same row
    const int P = max(1, min(32, (int)(gridDim.x / 16)));
    const int lines = (int)(row_stride / 8);
    const int lpw = (lines + P - 1) / P;
    const int line0 = ((b >> 4) % P) * lpw;
    const int line_end = min(lines, line0 + lpw);
    // warm-up loads per thread per step: the whole half slice (at N = 2048,
    // m = 8: 4096 lines at l = 2, 6144 at l = 3) from two workgroups per row
    // and XCD on;
    // a workgroup warming alone takes four lines per thread (2048 lines, part
    // of the slice: measured faster than warming all of it)
    constexpr int PF = L <= 2 ? 8 : 6;
    const int npf = P == 1 ? 4 : PF;
    uint32_t warm = 0, pf[PF] = {};

    cplx *const xb_mine0 = ws.xbuf + (((size_t)pair * 2 + 0) * 2 + a) * L * M;
    cplx *const xb_other0 = ws.xbuf + (((size_t)pair * 2 + 0) * 2 + (1 - a)) * L * M;
    const size_t parity_stride = (size_t)2 * L * M;

#pragma unroll 1
    for (int g = 0; g < n_iter; g++) {
        const cplx *bsk_g = bsk + (size_t)g * step_stride;
        const int32_t *idp[1] = {indices + (ct * n_iter + g) * m};

        // ---- 1. digits of row a (dec_fft.cuh:19-29), group grp at levels grp, grp+2, ...
#pragma unroll 1
        for (int it = 0; it < (L + 1) / 2; it++) {
            const int lev = 2 * it + grp;
            const int levc = min(lev, L - 1);      // past the last level: a copy, stored aside
            const int sh = log2_q - (levc + 1) * log2_beta - shift;
            cplx v1[1][4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int lo = (int)(((acc_t)(x[r] + offs) >> sh) & dmask) - half_beta;
                const int hi = (int)(((acc_t)(x[r + 4] + offs) >> sh) & dmask) - half_beta;
                v1[0][r] = cmul({(double)lo, -(double)hi}, zeta_pow<N, FULL>(ztab, t + r * T));
            }
            forward_wave<N, FULL, 1>(v1, lev < L ? own + lev * M : oth, ztab, t);
        }

        // ---- 2. publish this row's spectra at the bins thread u multiplies
        // (bin pair (u, M-1-u)), double-buffered by step parity
        const int r = u;
        cplx *xb_mine = xb_mine0 + (g & 1) * parity_stride + 2 * r;
        const cplx *xb_other = xb_other0 + (g & 1) * parity_stride + 2 * r;
        cplx dm[L][2];
#pragma unroll
        for (int lev = 0; lev < L; lev++) {
            dm[lev][0] = own[lev * M + phys<N>(r)];
            dm[lev][1] = own[lev * M + phys<N>(M - 1 - r)];
            if (local) {
                xb_mine[lev * M] = dm[lev][0];
                xb_mine[lev * M + 1] = dm[lev][1];
            } else {
                unsigned long long *d = reinterpret_cast<unsigned long long *>(xb_mine + lev * M);
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    __hip_atomic_store(d + 2 * k, (unsigned long long)__double_as_longlong(dm[lev][k].re),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    __hip_atomic_store(d + 2 * k + 1,
                                       (unsigned long long)__double_as_longlong(dm[lev][k].im),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
        }
        // every wave's stores must have reached L2 (same XCD) or memory
        // before thread 0 publishes the flag: a workgroup-scope release only
        // waits for LDS, so wait for all of this wave's memory operations
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(&ws.flags[pair * 2 + a], (uint32_t)(g + 1), __ATOMIC_RELAXED,
                               local ? __HIP_MEMORY_SCOPE_AGENT : __HIP_MEMORY_SCOPE_SYSTEM);

        // ---- 3. key patterns of output row a (mul_bootstrap.cuh:42-67): the
        // key rows times the rotation roots, per level and input row
        cplx W[L][1][2][1][2];
#pragma unroll
        for (int lev = 0; lev < L; lev++) {
#pragma unroll
            for (int b2 = 0; b2 < 2; b2++)
                W[lev][0][b2][0][0] = W[lev][0][b2][0][1] = {0.0, 0.0};
            cplx bkA[1][2][2], bkB[1][2][2];
            load_pair<L, 1>(bkA, bsk_g, 2 * r, a, lev, 0, m, M);
#pragma unroll 1
            for (int p = 0; p < m; p += 2) {
                load_pair<L, 1>(bkB, bsk_g, 2 * r, a, lev, p + 1 < m ? p + 1 : p, m, M);
                mac_pattern<N, 1, 1, FULL>(W[lev], bkA, idp, p, r, ztab);
                if (p + 1 < m) {
                    load_pair<L, 1>(bkA, bsk_g, 2 * r, a, lev, p + 2 < m ? p + 2 : p + 1, m, M);
                    mac_pattern<N, 1, 1, FULL>(W[lev], bkB, idp, p + 1, r, ztab);
                }
            }
        }

        // ---- 4. the partner row's spectra, then the sum over levels and rows
        if (threadIdx.x == 0 && !broken) {
            const bool ok =
                local ? wait_at_least<__HIP_MEMORY_SCOPE_AGENT>(&ws.flags[pair * 2 + 1 - a],
                                                                (uint32_t)(g + 1), spin_limit)
                      : wait_at_least<__HIP_MEMORY_SCOPE_SYSTEM>(&ws.flags[pair * 2 + 1 - a],
                                                                 (uint32_t)(g + 1), spin_limit);
            broken = !ok;
            if (!ok)
                atomicOr(ws.status, 1u);
        }
        __syncthreads();
        cplx dp[L][2];
#pragma unroll
        for (int lev = 0; lev < L; lev++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const unsigned long long *sp =
                    reinterpret_cast<const unsigned long long *>(xb_other + lev * M + k);
                unsigned long long re, im;
                if (local) {   // the L2 is shared: bypass only the L1
                    re = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    im = __hip_atomic_load(sp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    re = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    im = __hip_atomic_load(sp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
                dp[lev][k] = {__longlong_as_double((long long)re), __longlong_as_double((long long)im)};
            }
        cplx prod[2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
        for (int lev = 0; lev < L; lev++) {
            cplx acc0 = {0.0, 0.0}, acc1 = {0.0, 0.0};
#pragma unroll
            for (int b2 = 0; b2 < 2; b2++) {
                const cplx d0 = b2 == a ? dm[lev][0] : dp[lev][0];
                const cplx d1 = b2 == a ? dm[lev][1] : dp[lev][1];
                acc0 = cfma(W[lev][0][b2][0][0], d0, acc0);
                acc1 = cfma(W[lev][0][b2][0][1], d1, acc1);
            }
            prod[0] = cadd(prod[0], acc0);
            prod[1] = cadd(prod[1], acc1);
        }

        // L2 warm-up of the next step's half slice.  Issued here because the
        // memory counter retires in order: until the next step's publish only
        // LDS and VALU work follows, so no wait on these loads stalls a step.
#pragma unroll
        for (int i = 0; i < PF; i++) {
            warm ^= pf[i];
            pf[i] = 0;
        }
        if (g + 1 < n_iter) {
            const uint32_t *nxt = reinterpret_cast<const uint32_t *>(
                bsk + (size_t)(g + 1) * step_stride + (size_t)a * row_stride);
#pragma unroll
            for (int i = 0; i < PF; i++) {
                const int j = line0 + (int)threadIdx.x + i * (int)blockDim.x;
                if (i < npf && j < line_end)
                    pf[i] = nxt[(size_t)j * 32];
            }
        }

        // ---- 5. inverse transform of row a, untwist, centred mod q, round
        // (mul_bootstrap.cuh:70-93); one copy of the products per group.  No
        // thread reads own/oth between the barrier above and these writes.
        {
            cplx *c0 = own, *c1 = oth;
            c0[phys<N>(r)] = prod[0];
            c0[phys<N>(M - 1 - r)] = prod[1];
            c1[phys<N>(r)] = prod[0];
            c1[phys<N>(M - 1 - r)] = prod[1];
        }
        __syncthreads();
        cplx w[1][4];
        inverse_wave<N, FULL, 1>(w, grp == 0 ? own : oth, ztab, t);
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const cplx z = cmul(w[0][rr], zeta_pow<N, FULL>(ztab, -(t + rr * T)));
            double lo = z.re * inv_m;
            double hi = -z.im * inv_m;
            if constexpr (T32) {
                lo = rint(lo);
                hi = rint(hi);
                lo = lo - q * floor(fma(lo, inv_q, 0.5));
                hi = hi - q * floor(fma(hi, inv_q, 0.5));
                if (ACC == ACC_HI && g == n_iter - 1 && pair < batch && grp == 0) {
                    int64_t *dst = out + (size_t)pair * 2 * N + (size_t)a * N;
                    dst[t + rr * T] = (int64_t)lo;
                    dst[t + (rr + 4) * T] = (int64_t)hi;
                }
                x[rr] = (acc_t)((uint64_t)(int64_t)lo >> shift);
                x[rr + 4] = (acc_t)((uint64_t)(int64_t)hi >> shift);
            } else {
                lo = lo - q * rint(lo * inv_q);
                hi = hi - q * rint(hi * inv_q);
                lo = rint(lo);
                hi = rint(hi);
                if (log2_q >= 64) {
                    lo = lo >= 0x1p63 ? -0x1p63 : lo;
                    hi = hi >= 0x1p63 ? -0x1p63 : hi;
                }
                x[rr] = (acc_t)(int64_t)lo;
                x[rr + 4] = (acc_t)(int64_t)hi;
            }
        }
        // no barrier here: the next step's forward transforms write own/oth
        // only after their own first barrier, which every thread reaches
        // after reading its inverse input
    }

#pragma unroll
    for (int i = 0; i < PF; i++)
        warm ^= pf[i];
    if (warm == 0x9E3779B9u && ks.count < 0)   // ks.count >= 1 always: keeps the loads
        out[0] = (int64_t)warm;
    if (ACC != ACC_HI && pair < batch && grp == 0) {
        int64_t *dst = out + (size_t)pair * 2 * N + (size_t)a * N;
        const int sh_out = 64 - log2_q;
#pragma unroll
        for (int i = 0; i < 8; i++)
            dst[t + i * T] = T32 ? ((int64_t)((uint64_t)x[i] << sh_out) >> sh_out) : (int64_t)x[i];
    }
}

template <int N, int L, int MP>
static void launch_pair_acc(hipStream_t s, const int64_t *ct, int64_t st, const cplx *bk,
                            const int32_t *idx, int64_t *out, int64_t batch, int n_iter, int m,
                            int lb, int lq, KeySlots ks, PairWs ws, int spin)
{
    const dim3 grid((unsigned)(2 * pair_count(batch))), block(N / 4);
    if (lq <= 32)
        hipLaunchKernelGGL((br_pair_kernel<N, L, ACC_32, MP>), grid, block, 0, s, ct, st, bk, idx,
                           out, batch, n_iter, m, lb, lq, ks, ws, spin);
    else if (L * lb + 1 <= 32)
        hipLaunchKernelGGL((br_pair_kernel<N, L, ACC_HI, MP>), grid, block, 0, s, ct, st, bk, idx,
                           out, batch, n_iter, m, lb, lq, ks, ws, spin);
    else
        hipLaunchKernelGGL((br_pair_kernel<N, L, ACC_64, MP>), grid, block, 0, s, ct, st, bk, idx,
                           out, batch, n_iter, m, lb, lq, ks, ws, spin);
}

template <int N, int L>
static void launch_pair_l(hipStream_t s, const int64_t *ct, int64_t st, const cplx *bk,
                          const int32_t *idx, int64_t *out, int64_t batch, int n_iter, int m, int lb,
                          int lq, KeySlots ks, PairWs ws, int spin)
{
    if (m == 8)
        launch_pair_acc<N, L, 8>(s, ct, st, bk, idx, out, batch, n_iter, m, lb, lq, ks, ws, spin);
    else
        launch_pair_acc<N, L, 0>(s, ct, st, bk, idx, out, batch, n_iter, m, lb, lq, ks, ws, spin);
}

template <int N>
static void launch_pair_n(hipStream_t s, const int64_t *ct, int64_t st, const cplx *bk,
                          const int32_t *idx, int64_t *out, int64_t batch, int n_iter, int m, int lb,
                          int lq, int l, KeySlots ks, PairWs ws, int spin)
{
    switch (l) {
    case 1: launch_pair_l<N, 1>(s, ct, st, bk, idx, out, batch, n_iter, m, lb, lq, ks, ws, spin); break;
    case 2: launch_pair_l<N, 2>(s, ct, st, bk, idx, out, batch, n_iter, m, lb, lq, ks, ws, spin); break;
    default: launch_pair_l<N, 3>(s, ct, st, bk, idx, out, batch, n_iter, m, lb, lq, ks, ws, spin); break;
    }
}

// Ciphertexts the device holds with every workgroup resident: two
// workgroups per ciphertext; at N = 2048 one per CU (8 waves, up to 160 KiB
// of LDS), at N = 1024 two per CU (4 waves and <= 80 KiB each, within the
// two waves per SIMD the launch bounds give).
static int max_pairs_resident(int64_t degree)
{
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return 0;
    if (!cus[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        cus[dev] = n;
    }
    return degree == 1024 ? cus[dev] : cus[dev] / 2;
}

static int run_pair(void *stream, const int64_t *ct_init, int64_t ct_elems, const double *bsk,
                    const int64_t *bsk_dims, const double *all_rot_fft, int64_t rot_rows,
                    int64_t rot_cols, const int32_t *indices, int64_t idx_elems, int64_t *output,
                    int64_t batch, int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree,
                    void *workspace, int64_t workspace_bytes, int64_t spin_limit, KeySlots ks)
{
    clear_error();
    int64_t st = 0;
    int rc = validate_blind_rotation(ct_elems, bsk, bsk_dims, all_rot_fft, rot_rows, rot_cols,
                                     indices, idx_elems, output, batch, log2_beta, log2_q, l,
                                     degree, ct_init, &st);
    if (rc != BSLSTM_OK)
        return rc;
    if (!pair_degree(degree))
        return fail(BSLSTM_INVALID_ARGUMENT, "the pair form runs at degree 1024 or 2048");
    if (l > 3)
        return fail(BSLSTM_INVALID_ARGUMENT, "the pair form holds l <= 3");
    if (spin_limit < 1 || spin_limit > (1LL << 30))
        return fail(BSLSTM_INVALID_ARGUMENT, "spin_limit out of range [1, 2^30]");
    if (batch == 0)
        return BSLSTM_OK;
    if (!workspace || workspace_bytes < pair_ws_bytes(batch, l, degree))
        return fail(BSLSTM_INVALID_ARGUMENT, "workspace of %lld bytes; need %lld",
                    (long long)workspace_bytes, (long long)pair_ws_bytes(batch, l, degree));
    const int maxp = max_pairs_resident(degree);
    if (pair_count(batch) > maxp)
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "the pair form holds at most %d ciphertexts on this device (batch %lld)", maxp,
                    (long long)batch);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const PairWs ws = pair_ws(workspace, batch);
    const int64_t words = 1 + 4 * pair_count(batch);
    hipLaunchKernelGGL(pair_clear_kernel, dim3(4), dim3(256), 0, s, ws.status, words);
    const cplx *bk = reinterpret_cast<const cplx *>(bsk);
    const int ni = (int)bsk_dims[0], mm = (int)bsk_dims[4], lb = (int)log2_beta, lq = (int)log2_q;
    const int spin = (int)spin_limit;
    if (degree == 1024)
        launch_pair_n<1024>(s, ct_init, st, bk, indices, output, batch, ni, mm, lb, lq, (int)l, ks,
                            ws, spin);
    else
        launch_pair_n<2048>(s, ct_init, st, bk, indices, output, batch, ni, mm, lb, lq, (int)l, ks,
                            ws, spin);
    return launch_status("pair blind rotation");
}

}  // namespace bslstm

extern "C" int64_t bslstm_mul_bootstrap_pair_workspace_bytes(int64_t batch, int64_t l,
                                                             int64_t degree)
{
    if (batch < 0 || l < 1 || l > 3 || !bslstm::pair_degree(degree))
        return -1;
    return bslstm::pair_ws_bytes(batch, l, degree);
}

extern "C" int bslstm_mul_bootstrap_pair(void *stream, const int64_t *ct_init, int64_t ct_elems,
                                         const double *bsk, const int64_t bsk_dims[6],
                                         const double *all_rot_fft, int64_t rot_rows,
                                         int64_t rot_cols, const int32_t *indices,
                                         int64_t idx_elems, int64_t *output, int64_t batch,
                                         int64_t log2_beta, int64_t log2_q, int64_t l,
                                         int64_t degree, void *workspace, int64_t workspace_bytes,
                                         int64_t spin_limit)
{
    return bslstm::run_pair(stream, ct_init, ct_elems, bsk, bsk_dims, all_rot_fft, rot_rows,
                            rot_cols, indices, idx_elems, output, batch, log2_beta, log2_q, l,
                            degree, workspace, workspace_bytes, spin_limit,
                            bslstm::KeySlots{nullptr, 0, 1});
}

extern "C" int bslstm_mul_bootstrap_pair_slots(void *stream, const int64_t *ct_init,
                                               int64_t ct_elems, const double *bsk_pool,
                                               const int64_t bsk_dims[6], const int32_t *slots,
                                               int64_t slot_count, const int32_t *indices,
                                               int64_t idx_elems, int64_t *output, int64_t batch,
                                               int64_t log2_beta, int64_t log2_q, int64_t l,
                                               int64_t degree, void *workspace,
                                               int64_t workspace_bytes, int64_t spin_limit)
{
    if (!slots || slot_count < 1 || slot_count > 0x7fffffffLL || !bsk_dims)
        return bslstm::fail(BSLSTM_INVALID_ARGUMENT, "slots and 1 <= slot_count required");
    int64_t stride = 1;
    for (int i = 0; i < 6; i++)
        stride *= bsk_dims[i];
    return bslstm::run_pair(stream, ct_init, ct_elems, bsk_pool, bsk_dims, nullptr, 0, 0, indices,
                            idx_elems, output, batch, log2_beta, log2_q, l, degree, workspace,
                            workspace_bytes, spin_limit,
                            bslstm::KeySlots{slots, stride, (int32_t)slot_count});
}
