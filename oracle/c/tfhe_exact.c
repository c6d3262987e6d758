/*
 * Exact-integer oracle for the collapsed blind rotation.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ (ctypes) as the checker.  The
 * product library never links it.
 *
 * It computes what csrc/mul_bootstrap.cuh:20-136 of the reference computes,
 * but with exact integer arithmetic in place of the FP64 FFT: every
 * negacyclic product is a schoolbook sum accumulated in signed 128-bit
 * integers, then reduced to the centred representative mod q = 2^log_q.
 *
 *   per step g (mul_bootstrap.cuh:122-131):
 *     dec[bc] = signed digits of ct by the offset rule (dec_fft.cuh:4-31),
 *               bc = 2*level + component
 *     for a in {0,1}:
 *       P_bc  = sum_p X^{idx[g,p]} * BSK[g,a,bc,p]        (mul_bootstrap.cuh:51-59
 *                                                          in the coefficient domain)
 *       ct[a] = sum_bc P_bc * dec[bc]  mod q               (:61-92)
 *
 * With exact arithmetic the result is the ideal one the FP64 path rounds to;
 * tests report the FP64 deviation from it (phase-error std) and require the
 * decrypted messages to agree.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
/* sums wrap mod 2^128 (well defined for an unsigned type); every result is
   read mod 2^log_q <= 2^64, which that wrap preserves, and the blind
   rotation's sums (small digits times key entries) never wrap at all */
typedef unsigned __int128 u128;

static int64_t centred(u128 v, int log_q)
{
    uint64_t r = (uint64_t)v; /* v mod 2^64 */
    if (log_q >= 64)
        return (int64_t)r;
    uint64_t q = 1ULL << log_q;
    r &= q - 1;
    if (r >= (q >> 1))
        return (int64_t)(r - q);
    return (int64_t)r;
}

/* dec_fft.cuh:4-31: digits (L, 2N) in [-beta/2, beta/2) */
static void decompose(const int64_t *ct, int64_t *dec, int N, int L, int log_beta, int log_q)
{
    uint64_t off = 0;
    for (int i = 1; i <= L; i++)
        off += 1ULL << (log_q - i * log_beta + log_beta - 1);
    int rs = log_q - L * log_beta - 1;
    if (rs >= 0 && rs < 64)
        off += 1ULL << rs;
    uint64_t mask = (1ULL << log_beta) - 1;
    for (int k = 0; k < 2 * N; k++) {
        uint64_t st = (uint64_t)ct[k] + off;
        for (int i = 1; i <= L; i++) {
            int sh = log_q - i * log_beta;
            dec[(size_t)(i - 1) * 2 * N + k] =
                (int64_t)((st >> sh) & mask) - (int64_t)(1LL << (log_beta - 1));
        }
    }
}

/* out += X^e * p  (negacyclic, e in [0, 2N)) */
static void add_rotated(u128 *out, const int64_t *p, int e, int N)
{
    for (int k = 0; k < N; k++) {
        int t = k + e;
        int neg = 0;
        while (t >= N) {
            t -= N;
            neg ^= 1;
        }
        out[t] += neg ? (u128)0 - (u128)(i128)p[k] : (u128)(i128)p[k];
    }
}

/* acc += a * d  (negacyclic), a centred mod q, d small */
static void negacyclic_mac(u128 *acc, const int64_t *a, const int64_t *d, int N)
{
    for (int i = 0; i < N; i++) {
        i128 di = d[i];
        if (di == 0)
            continue;
        for (int j = 0; j < N - i; j++)
            acc[i + j] += (u128)(di * a[j]);   /* |product| < 2^126: exact */
        for (int j = N - i; j < N; j++)
            acc[i + j - N] -= (u128)(di * a[j]);
    }
}

/* Exact negacyclic product mod 2^log_q (centred). */
void tfhe_exact_negacyclic_mul(const int64_t *a, const int64_t *b, int64_t *out, int N, int log_q)
{
    u128 *acc = (u128 *)calloc((size_t)N, sizeof(u128));
    negacyclic_mac(acc, a, b, N);
    for (int k = 0; k < N; k++)
        out[k] = centred(acc[k], log_q);
    free(acc);
}

/*
 * ct_in  (B, 2, N) int64;  stride_in = 2N or 0 (broadcast, mul_bootstrap.cuh:162-164)
 * bsk    (n_iter, 2[a], L, 2[L/R], m, N) int64 coefficient-domain key
 * idx    (B, n_iter, m) int32
 * out    (B, 2, N) int64
 */
void tfhe_exact_blind_rotate(const int64_t *ct_in, long stride_in, const int64_t *bsk,
                             const int32_t *idx, int64_t *out, int B, int n_iter, int m,
                             int N, int L, int log_beta, int log_q)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < B; b++) {
        int64_t *ct = (int64_t *)malloc(sizeof(int64_t) * 2 * N);
        int64_t *dec = (int64_t *)malloc(sizeof(int64_t) * (size_t)L * 2 * N);
        int64_t *P = (int64_t *)malloc(sizeof(int64_t) * N);
        u128 *rot = (u128 *)malloc(sizeof(u128) * N);
        u128 *acc = (u128 *)malloc(sizeof(u128) * 2 * N);
        memcpy(ct, ct_in + (size_t)b * stride_in, sizeof(int64_t) * 2 * N);
        const int BC = 2 * L;
        for (int g = 0; g < n_iter; g++) {
            const int32_t *id = idx + ((size_t)b * n_iter + g) * m;
            decompose(ct, dec, N, L, log_beta, log_q);
            memset(acc, 0, sizeof(u128) * 2 * N);
            for (int a = 0; a < 2; a++) {
                for (int bc = 0; bc < BC; bc++) {
                    memset(rot, 0, sizeof(u128) * N);
                    for (int p = 0; p < m; p++) {
                        const int64_t *key =
                            bsk + ((((size_t)g * 2 + a) * BC + bc) * m + p) * (size_t)N;
                        add_rotated(rot, key, id[p], N);
                    }
                    for (int k = 0; k < N; k++)
                        P[k] = centred(rot[k], log_q);
                    /* dec layout: level-major, then component: bc = 2*lev + comp */
                    negacyclic_mac(acc + (size_t)a * N, P, dec + (size_t)bc * N, N);
                }
            }
            for (int k = 0; k < 2 * N; k++)
                ct[k] = centred(acc[k], log_q);
        }
        memcpy(out + (size_t)b * 2 * N, ct, sizeof(int64_t) * 2 * N);
        free(ct);
        free(dec);
        free(P);
        free(rot);
        free(acc);
    }
}
