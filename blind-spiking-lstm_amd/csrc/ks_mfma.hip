// LWE key switch as exact int8 MFMA GEMMs (gfx950 v_mfma_i32_32x32x32_i8).
//
// The key switch of schemas/LWE_jax.py:97-117 is the product of the digit
// matrix D (batch x K, K = n_in*l, signed digits of the float rule of
// polynomial_jax.py:90-111, |d| <= beta/2 <= 64) with the key matrix
// (K x (n_out+1), the body key as the last column), mod 2^64.  The key is cut
// into eight balanced signed bytes, key = sum_b k_b 2^(8b) mod 2^64 with
// k_b in [-128, 127]; every limb product D x k_b is an int8 GEMM whose int32
// sums are exact (|sum| <= K * 64 * 128 < 2^31 for K < 2^18), and
// sum_b (int64)acc_b << 8b mod 2^64 is the exact product.  This is the
// "split" form: integer-exact, on the matrix cores, at any modulus <= 2^64.
//
// Layouts (chosen for 16-byte loads per lane, one 1 KiB block per wave and
// MFMA step):
//   k = i*l + lev (input-major, level-minor)
//   key tiles  [kt][nt][b][lane*16 + j]: limb b of key row k = kt*32 +
//              (lane/32)*16 + j, column n = nt*32 + lane%32
//   digit tiles [mt][kt][lane*16 + j]:  digit of row m = mt*32 + lane%32,
//              k = kt*32 + (lane/32)*16 + j
// Lane l of both MFMA operands carries the same 16 k values, so the product
// does not depend on the instruction's internal k order.  The 32x32 int32
// result sits in the usual layout: register r of lane l is row
// 8*(r/4) + 4*(l/32) + r%4, column l%32.
//
// The reduction over K is split across gridDim.z workgroups when the output
// alone gives too few; partial int64 results meet in the outputs through
// exact 64-bit integer atomics (order-independent, bit-reproducible).
#include <algorithm>
#include <mutex>

#include "capi_internal.h"
#include "torus_device.h"

namespace bslstm {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int KM_WAVES = 4;   // waves per workgroup, one 32-column tile each

__global__ void ks_pack_key_kernel(const int64_t *__restrict__ ksk_a,
                                   const int64_t *__restrict__ ksk_b, int l, int n_in, int n_out,
                                   int KT, int NT, int8_t *__restrict__ tiles)
{
    // one thread per (kt, nt, lane): 16 k values x 8 limbs
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)KT * NT * 64;
    if (gid >= total)
        return;
    const int lane = (int)(gid % 64);
    const int64_t blk = gid / 64;
    const int nt = (int)(blk % NT);
    const int kt = (int)(blk / NT);
    const int n = nt * 32 + lane % 32;
    const int K = n_in * l;
    int8_t limbs[8][16];
    for (int j = 0; j < 16; j++) {
        const int k = kt * 32 + (lane / 32) * 16 + j;
        uint64_t v = 0;
        if (k < K && n <= n_out) {
            const int i = k / l, lev = k % l;
            v = (uint64_t)(n < n_out ? ksk_a[((size_t)lev * n_in + i) * n_out + n]
                                     : ksk_b[(size_t)lev * n_in + i]);
        }
        int carry = 0;
        for (int b = 0; b < 8; b++) {
            int d = (int)((v >> (8 * b)) & 255) + carry;
            carry = d >= 128;
            limbs[b][j] = (int8_t)(d - carry * 256);
        }
    }
    for (int b = 0; b < 8; b++) {
        v4i w;
        for (int q = 0; q < 4; q++)
            w[q] = (int)((uint32_t)(uint8_t)limbs[b][4 * q] | ((uint32_t)(uint8_t)limbs[b][4 * q + 1] << 8) |
                         ((uint32_t)(uint8_t)limbs[b][4 * q + 2] << 16) |
                         ((uint32_t)(uint8_t)limbs[b][4 * q + 3] << 24));
        *reinterpret_cast<v4i *>(tiles + (((size_t)kt * NT + nt) * 8 + b) * 1024 + lane * 16) = w;
    }
}

// One workgroup per (row tile, run of KD_TILES k tiles): the digits of its 32
// rows x 32*KD_TILES k are computed into LDS by one thread per (row, input),
// consecutive threads on consecutive inputs of a row, each entering the
// digit recursion at its first level in the run; then every thread writes
// one 16-byte lane block of a tile.
constexpr int KD_TILES = 4;
constexpr int KD_K = 32 * KD_TILES;

__global__ __launch_bounds__(256) void ks_digits_kernel(const int64_t *__restrict__ a_in, int batch,
                                                        int n_in, int l, int lb, int lq, int KT,
                                                        int MT, int8_t *__restrict__ tiles)
{
    __shared__ int8_t dg[32][KD_K + 16];
    const int mt = blockIdx.x;
    const int kt0 = blockIdx.y * KD_TILES;
    const int kbase = kt0 * 32;
    const int K = n_in * l;
    for (int e = threadIdx.x; e < 32 * (KD_K + 16) / 4; e += 256)
        reinterpret_cast<int32_t *>(&dg[0][0])[e] = 0;
    __syncthreads();
    const int i_lo = kbase / l;
    const int i_hi = min(n_in - 1, (kbase + KD_K - 1) / l);
    const int ni = i_hi - i_lo + 1;
    for (int e = threadIdx.x; e < 32 * ni; e += 256) {
        const int rr = e / ni, i = i_lo + (e - rr * ni);
        const int row = mt * 32 + rr;
        if (row >= batch)
            continue;
        const int k_first = max(i * l, kbase);
        const int k_end = min(min(i * l + l, kbase + KD_K), K);
        int64_t r = residual_at(tmod((uint64_t)a_in[(size_t)row * n_in + i], lq),
                                lq - (k_first - i * l) * lb, lq);
        for (int k = k_first; k < k_end; k++)
            dg[rr][k - kbase] = (int8_t)next_digit(r, lq - (k - i * l + 1) * lb);
    }
    __syncthreads();
    const int tile = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (kt0 + tile >= KT)
        return;
    const int8_t *src = &dg[lane & 31][tile * 32 + (lane >> 5) * 16];
    v4i w;
#pragma unroll
    for (int q = 0; q < 4; q++)
        w[q] = (int)((uint32_t)(uint8_t)src[4 * q] | ((uint32_t)(uint8_t)src[4 * q + 1] << 8) |
                     ((uint32_t)(uint8_t)src[4 * q + 2] << 16) | ((uint32_t)(uint8_t)src[4 * q + 3] << 24));
    *reinterpret_cast<v4i *>(tiles + ((size_t)mt * KT + kt0 + tile) * 1024 + lane * 16) = w;
}

__global__ void ks_out_init_kernel(const int64_t *__restrict__ b_in, int batch, int n_out,
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

__global__ void ks_out_final_kernel(int batch, int n_out, int lq, int64_t *__restrict__ a_out,
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

// tile_slots (optional): the key of row tile mt is ktiles + tile_slots[mt] *
// slot_bytes (clamped to [0, slot_count)), for batches that mix clients
__global__ __launch_bounds__(64 * KM_WAVES, 2) void ks_mfma_kernel(
    const int8_t *__restrict__ dtiles, const int8_t *__restrict__ ktiles, int batch, int n_out,
    int KT, int NT, int kt_per_split, const int32_t *__restrict__ tile_slots, int64_t slot_bytes,
    int slot_count, int64_t *__restrict__ a_out, int64_t *__restrict__ b_out)
{
    const int lane = threadIdx.x & 63;
    const int nt = blockIdx.x * KM_WAVES + (threadIdx.x >> 6);
    const int mt = blockIdx.y;
    if (tile_slots) {
        int sl = tile_slots[mt];
        sl = sl < 0 ? 0 : (sl >= slot_count ? slot_count - 1 : sl);
        ktiles += (size_t)sl * slot_bytes;
    }
    const int kt0 = blockIdx.z * kt_per_split;
    const int kt1 = min(KT, kt0 + kt_per_split);
    if (nt >= NT)
        return;
    v16i acc[8];
#pragma unroll
    for (int b = 0; b < 8; b++)
        acc[b] = (v16i){0};
    const int8_t *ap = dtiles + ((size_t)mt * KT) * 1024 + lane * 16;
    const int8_t *bp = ktiles + lane * 16;
#pragma unroll 1
    for (int kt = kt0; kt < kt1; kt++) {
        const v4i a = *reinterpret_cast<const v4i *>(ap + (size_t)kt * 1024);
        const int8_t *bk = bp + ((size_t)kt * NT + nt) * 8 * 1024;
        v4i bb[8];
#pragma unroll
        for (int b = 0; b < 8; b++)
            bb[b] = *reinterpret_cast<const v4i *>(bk + b * 1024);
#pragma unroll
        for (int b = 0; b < 8; b++)
            acc[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bb[b], acc[b], 0, 0, 0);
    }
    const int n = nt * 32 + (lane & 31);
    if (n > n_out)
        return;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = mt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m >= batch)
            continue;
        uint64_t v = 0;
#pragma unroll
        for (int b = 0; b < 8; b++)
            v += (uint64_t)(int64_t)acc[b][r] << (8 * b);
        unsigned long long *dst = reinterpret_cast<unsigned long long *>(
            n < n_out ? a_out + (size_t)m * n_out + n : b_out + m);
        atomicAdd(dst, (unsigned long long)v);
    }
}

// Batches of two or more row tiles with one key: a workgroup covers one
// 32-column tile and KM4_MT = 4 row tiles; wave w takes row tiles 2*(w/2),
// 2*(w/2)+1 and limbs 4*(w%2) .. 4*(w%2)+3, so each wave loads 2 digit tiles
// and 4 key tiles per k step for 8 MFMAs (the kernel above loads 9 for 8) and
// a key tile serves four row tiles.  Waves past the last row tile only join
// the barrier.  The two limb halves of an output meet in LDS before the
// atomic add.
constexpr int KM4_MT = 4;

__global__ __launch_bounds__(256, 2) void ks_mfma4_kernel(const int8_t *__restrict__ dtiles,
                                                       const int8_t *__restrict__ ktiles, int batch,
                                                       int n_out, int KT, int NT, int MT,
                                                       int kt_per_split, int64_t *__restrict__ a_out,
                                                       int64_t *__restrict__ b_out)
{
    __shared__ uint64_t upper[2][2][16][64];   // [row-tile pair][tile][register][lane]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mp = w >> 1, lh = w & 1;
    const int nt = blockIdx.x;
    const int mtA = blockIdx.y * KM4_MT + 2 * mp;
    const int kt0 = blockIdx.z * kt_per_split;
    const int kt1 = min(KT, kt0 + kt_per_split);
    v16i acc[2][4];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[t][j] = (v16i){0};
    if (mtA < MT) {   // wave-uniform; a missing second tile re-reads the first, its sums unused
        const int8_t *ap0 = dtiles + ((size_t)mtA * KT) * 1024 + lane * 16;
        const int8_t *ap1 = dtiles + ((size_t)min(mtA + 1, MT - 1) * KT) * 1024 + lane * 16;
        const int8_t *bp = ktiles + (size_t)(4 * lh) * 1024 + lane * 16;
#pragma unroll 1
        for (int kt = kt0; kt < kt1; kt++) {
            const v4i a0 = *reinterpret_cast<const v4i *>(ap0 + (size_t)kt * 1024);
            const v4i a1 = *reinterpret_cast<const v4i *>(ap1 + (size_t)kt * 1024);
            const int8_t *bk = bp + ((size_t)kt * NT + nt) * 8 * 1024;
            v4i bb[4];
#pragma unroll
            for (int j = 0; j < 4; j++)
                bb[j] = *reinterpret_cast<const v4i *>(bk + j * 1024);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[0][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, bb[j], acc[0][j], 0, 0, 0);
                acc[1][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, bb[j], acc[1][j], 0, 0, 0);
            }
        }
    }
    // this wave's limbs combined, sum_j acc_j << 8*(4*lh + j) mod 2^64, one
    // output at a time (no array of partial sums held beside the accumulators)
    auto limbs = [&](int t, int r) -> uint64_t {
        uint64_t v = 0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            v += (uint64_t)(int64_t)acc[t][j][r] << (8 * (4 * lh + j));
        return v;
    };
    if (lh == 1)
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                upper[mp][t][r][lane] = limbs(t, r);
    __syncthreads();
    const int n = nt * 32 + (lane & 31);
    if (lh != 0 || n > n_out)
        return;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int mt = mtA + t;
        if (mt >= MT)
            continue;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int m = mt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (m >= batch)
                continue;
            unsigned long long *dst = reinterpret_cast<unsigned long long *>(
                n < n_out ? a_out + (size_t)m * n_out + n : b_out + m);
            atomicAdd(dst, (unsigned long long)(limbs(t, r) + upper[mp][t][r][lane]));
        }
    }
}

static int tiles_k(int64_t n_in, int64_t l) { return (int)((n_in * l + 31) / 32); }
static int tiles_n(int64_t n_out) { return (int)((n_out + 1 + 31) / 32); }

}  // namespace bslstm

using namespace bslstm;

extern "C" int64_t bslstm_ks_mfma_key_bytes(int64_t l, int64_t n_in, int64_t n_out)
{
    if (l < 1 || n_in < 1 || n_out < 1 || n_in * l >= (1 << 18))
        return -1;
    return (int64_t)tiles_k(n_in, l) * tiles_n(n_out) * 8 * 1024;
}

extern "C" int bslstm_ks_mfma_prepare_key(void *stream, const int64_t *ksk_a, const int64_t *ksk_b,
                                          int64_t l, int64_t n_in, int64_t n_out, int8_t *key_tiles)
{
    clear_error();
    if (bslstm_ks_mfma_key_bytes(l, n_in, n_out) < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad l/n_in/n_out (n_in*l < 2^18)");
    if (!ksk_a || !ksk_b || !key_tiles)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL operand");
    const int KT = tiles_k(n_in, l), NT = tiles_n(n_out);
    const int64_t total = (int64_t)KT * NT * 64;
    hipLaunchKernelGGL(ks_pack_key_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, ksk_a, ksk_b, (int)l, (int)n_in, (int)n_out, KT, NT,
                       key_tiles);
    return launch_status("key-switch key tiles");
}

extern "C" int64_t bslstm_ks_mfma_workspace_bytes(int64_t batch, int64_t l, int64_t n_in)
{
    if (batch < 0 || l < 1 || n_in < 1)
        return -1;
    return ((batch + 31) / 32) * (int64_t)tiles_k(n_in, l) * 1024;
}

static int ks_mfma_run(void *stream, const int64_t *a_in, const int64_t *b_in,
                       const int8_t *key_tiles, const int32_t *tile_slots, int64_t slot_count,
                       int64_t batch, int64_t n_in, int64_t n_out, int64_t log2_beta_ks,
                       int64_t l_ks, int64_t log2_q, void *workspace, int64_t workspace_bytes,
                       int64_t *a_out, int64_t *b_out)
{
    clear_error();
    if (log2_q < 2 || log2_q > 64)
        return fail(BSLSTM_INVALID_ARGUMENT, "log2_q must be in [2, 64]");
    if (l_ks < 1 || log2_beta_ks < 1 || log2_beta_ks > 7 || l_ks * log2_beta_ks > log2_q)
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "int8 digits need 1 <= log2_beta_ks <= 7 and l_ks*log2_beta_ks <= log2_q");
    // row tiles index the grid's y dimension (at most 65535)
    if (bslstm_ks_mfma_key_bytes(l_ks, n_in, n_out) < 0 || batch < 0 || batch > 65535LL * 32)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad shapes (n_in*l_ks < 2^18, batch <= 2097120)");
    if (batch == 0)
        return BSLSTM_OK;
    if (workspace_bytes < bslstm_ks_mfma_workspace_bytes(batch, l_ks, n_in))
        return fail(BSLSTM_INVALID_ARGUMENT, "workspace too small");
    if (!a_in || !b_in || !key_tiles || !workspace || !a_out || !b_out || b_out == b_in ||
        a_out == a_in)
        return fail(BSLSTM_INVALID_ARGUMENT, "NULL or aliased operand");
    hipStream_t s = (hipStream_t)stream;
    const int KT = tiles_k(n_in, l_ks), NT = tiles_n(n_out), MT = (int)((batch + 31) / 32);
    int8_t *dt = reinterpret_cast<int8_t *>(workspace);
    hipLaunchKernelGGL(ks_digits_kernel, dim3((unsigned)MT, (unsigned)((KT + KD_TILES - 1) / KD_TILES)),
                       dim3(256), 0, s, a_in, (int)batch, (int)n_in, (int)l_ks, (int)log2_beta_ks,
                       (int)log2_q, KT, MT, dt);
    const int64_t total = batch * (n_out + 1);
    const unsigned eb = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(ks_out_init_kernel, dim3(eb), dim3(256), 0, s, b_in, (int)batch, (int)n_out,
                       a_out, b_out);
    if (!tile_slots && MT >= 2) {
        // one key, two or more row tiles: four row tiles per workgroup (the
        // key is then read once per four row tiles, not once per tile);
        // split K until the grid has ~4 workgroups per CU
        const int gy = (MT + KM4_MT - 1) / KM4_MT;
        int ks = (int)std::max<int64_t>(1, std::min<int64_t>(KT, 1024 / ((int64_t)NT * gy)));
        const int kps = (KT + ks - 1) / ks;
        ks = (KT + kps - 1) / kps;
        hipLaunchKernelGGL(ks_mfma4_kernel, dim3(NT, gy, ks), dim3(256), 0, s, dt, key_tiles,
                           (int)batch, (int)n_out, KT, NT, MT, kps, a_out, b_out);
    } else {
        const int gx = (NT + KM_WAVES - 1) / KM_WAVES;
        // split K until the grid has ~2 workgroups per CU
        int ks = (int)std::max<int64_t>(1, std::min<int64_t>(KT, 512 / ((int64_t)gx * MT)));
        const int kps = (KT + ks - 1) / ks;
        ks = (KT + kps - 1) / kps;
        hipLaunchKernelGGL(ks_mfma_kernel, dim3(gx, MT, ks), dim3(64 * KM_WAVES), 0, s, dt,
                           key_tiles, (int)batch, (int)n_out, KT, NT, kps, tile_slots,
                           bslstm_ks_mfma_key_bytes(l_ks, n_in, n_out), (int)slot_count, a_out,
                           b_out);
    }
    if (int st = launch_status("key switch (MFMA)"))
        return st;
    if (log2_q < 64)
        hipLaunchKernelGGL(ks_out_final_kernel, dim3(eb), dim3(256), 0, s, (int)batch, (int)n_out,
                           (int)log2_q, a_out, b_out);
    return launch_status("key switch (MFMA)");
}

extern "C" int bslstm_key_switch_lwe_mfma(void *stream, const int64_t *a_in, const int64_t *b_in,
                                          const int8_t *key_tiles, int64_t batch, int64_t n_in,
                                          int64_t n_out, int64_t log2_beta_ks, int64_t l_ks,
                                          int64_t log2_q, void *workspace, int64_t workspace_bytes,
                                          int64_t *a_out, int64_t *b_out)
{
    return ks_mfma_run(stream, a_in, b_in, key_tiles, nullptr, 1, batch, n_in, n_out, log2_beta_ks,
                       l_ks, log2_q, workspace, workspace_bytes, a_out, b_out);
}

extern "C" int bslstm_key_switch_lwe_mfma_slots(void *stream, const int64_t *a_in,
                                                const int64_t *b_in, const int8_t *key_pool,
                                                const int32_t *tile_slots, int64_t slot_count,
                                                int64_t batch, int64_t n_in, int64_t n_out,
                                                int64_t log2_beta_ks, int64_t l_ks, int64_t log2_q,
                                                void *workspace, int64_t workspace_bytes,
                                                int64_t *a_out, int64_t *b_out)
{
    if (!tile_slots || slot_count < 1 || slot_count > 0x7fffffffLL)
        return fail(BSLSTM_INVALID_ARGUMENT, "tile_slots and 1 <= slot_count required");
    return ks_mfma_run(stream, a_in, b_in, key_pool, tile_slots, slot_count, batch, n_in, n_out,
                       log2_beta_ks, l_ks, log2_q, workspace, workspace_bytes, a_out, b_out);
}

// ---------------------------------------------------------------------------
// Allocation-free form.  The digit tiles (the only device memory the key
// switch needs beyond its operands) live in a file-scope device array of
// KS_NOALLOC_SCRATCH_BYTES; every GPU holds its own copy of it until the
// process ends.  Host code never passes the array's name to a kernel: its
// device address is looked up once per device with hipGetSymbolAddress (that
// device current) and kept.  All calls on one device share the array, so
// calls that overlap on one device (two streams, or two host threads) race;
// the digit kernel rewrites every tile a call reads, so nothing is cleared.
// ---------------------------------------------------------------------------
namespace bslstm {

constexpr int64_t KS_NOALLOC_SCRATCH_BYTES = 16LL << 20;
constexpr int KS_NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) int8_t scratch[KS_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[KS_NOALLOC_MAX_DEVICES];

// Device address of `scratch` on the device that runs `stream`.
static int scratch_address(hipStream_t stream, void **out)
{
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess)
        return fail(BSLSTM_INTERNAL, "hipGetDevice: %s", hipGetErrorString(e));
    int dev = cur;
    if (stream) {
        hipDevice_t sd = 0;
        e = hipStreamGetDevice(stream, &sd);
        if (e != hipSuccess)
            return fail(BSLSTM_INVALID_ARGUMENT, "hipStreamGetDevice: %s", hipGetErrorString(e));
        dev = (int)sd;
    }
    if (dev < 0 || dev >= KS_NOALLOC_MAX_DEVICES)
        return fail(BSLSTM_INVALID_ARGUMENT, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        if (dev != cur && (e = hipSetDevice(dev)) != hipSuccess)
            return fail(BSLSTM_INTERNAL, "hipSetDevice(%d): %s", dev, hipGetErrorString(e));
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (dev != cur)
            (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(BSLSTM_INTERNAL, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BSLSTM_OK;
}

}  // namespace bslstm

extern "C" int64_t bslstm_key_switch_lwe_mfma_noalloc_max_bytes(void)
{
    return bslstm::KS_NOALLOC_SCRATCH_BYTES;
}

extern "C" int bslstm_key_switch_lwe_mfma_noalloc(void *stream, const int64_t *a_in,
                                                  const int64_t *b_in, const int8_t *key_tiles,
                                                  int64_t batch, int64_t n_in, int64_t n_out,
                                                  int64_t log2_beta_ks, int64_t l_ks,
                                                  int64_t log2_q, int64_t *a_out, int64_t *b_out)
{
    clear_error();
    const int64_t need = bslstm_ks_mfma_workspace_bytes(batch, l_ks, n_in);
    if (need < 0)
        return fail(BSLSTM_INVALID_ARGUMENT, "bad shapes (batch >= 0, l_ks >= 1, n_in >= 1)");
    if (need > KS_NOALLOC_SCRATCH_BYTES)
        return fail(BSLSTM_INVALID_ARGUMENT,
                    "batch %lld x n_in %lld x l_ks %lld needs %lld bytes of digit tiles; the "
                    "allocation-free key switch holds at most %lld",
                    (long long)batch, (long long)n_in, (long long)l_ks, (long long)need,
                    (long long)KS_NOALLOC_SCRATCH_BYTES);
    if (batch == 0)
        return ks_mfma_run(stream, a_in, b_in, key_tiles, nullptr, 1, batch, n_in, n_out,
                           log2_beta_ks, l_ks, log2_q, nullptr, 0, a_out, b_out);
    void *ws = nullptr;
    if (int st = scratch_address((hipStream_t)stream, &ws))
        return st;
    return ks_mfma_run(stream, a_in, b_in, key_tiles, nullptr, 1, batch, n_in, n_out, log2_beta_ks,
                       l_ks, log2_q, ws, KS_NOALLOC_SCRATCH_BYTES, a_out, b_out);
}
