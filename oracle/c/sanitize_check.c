/*
 * Host-side check of the exact C oracle under AddressSanitizer and
 * UndefinedBehaviorSanitizer (built and run by `make -C oracle sanitize`).
 * TEST INFRASTRUCTURE ONLY.
 *
 *  1. tfhe_exact_negacyclic_mul equals a schoolbook product computed here
 *     with exact 128-bit sums, for several N and log_q.
 *  2. tfhe_exact_blind_rotate on a random key and exponents, over several
 *     batch shapes: a broadcast test polynomial (stride 0) gives the same
 *     outputs as the same polynomial repeated per ciphertext (stride 2N); a
 *     zero key gives the zero ciphertext.
 * Any out-of-bounds access, overflow of a signed type or misaligned access
 * in the oracle stops the run with a sanitizer report (non-zero exit).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void tfhe_exact_negacyclic_mul(const int64_t *a, const int64_t *b, int64_t *out, int N, int log_q);
void tfhe_exact_blind_rotate(const int64_t *ct_in, long stride_in, const int64_t *bsk,
                             const int32_t *idx, int64_t *out, int B, int n_iter, int m,
                             int N, int L, int log_beta, int log_q);

static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64(void)
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}

/* centred representative of v mod 2^log_q */
static int64_t centred_u(uint64_t v, int log_q)
{
    const int s = 64 - log_q;
    return (int64_t)(v << s) >> s;
}

static int64_t rand_class(int log_q) { return centred_u(next_u64(), log_q); }

static int check_mul(int N, int log_q)
{
    int64_t *a = malloc(sizeof(int64_t) * N), *b = malloc(sizeof(int64_t) * N);
    int64_t *got = malloc(sizeof(int64_t) * N);
    for (int i = 0; i < N; i++) {
        a[i] = rand_class(log_q);
        b[i] = rand_class(log_q);
    }
    tfhe_exact_negacyclic_mul(a, b, got, N, log_q);
    int bad = 0;
    for (int k = 0; k < N; k++) {
        uint64_t acc = 0;   /* the class mod 2^64 suffices for log_q <= 64 */
        for (int i = 0; i < N; i++) {
            const int j = k - i;
            const uint64_t p = (uint64_t)a[i] * (uint64_t)b[j >= 0 ? j : j + N];
            acc += j >= 0 ? p : 0 - p;
        }
        if (centred_u(acc, log_q) != got[k])
            bad++;
    }
    free(a);
    free(b);
    free(got);
    if (bad)
        printf("negacyclic_mul N=%d log_q=%d: %d wrong coefficients\n", N, log_q, bad);
    return bad;
}

static int check_blind_rotate(int B, int n_iter, int m, int N, int L, int log_beta, int log_q)
{
    const size_t key_n = (size_t)n_iter * 2 * L * 2 * m * N;
    int64_t *bsk = malloc(sizeof(int64_t) * key_n);
    int64_t *zero = calloc(key_n, sizeof(int64_t));
    int32_t *idx = malloc(sizeof(int32_t) * (size_t)B * n_iter * m);
    int64_t *lut = malloc(sizeof(int64_t) * 2 * N);
    int64_t *lut_rep = malloc(sizeof(int64_t) * (size_t)B * 2 * N);
    int64_t *o1 = malloc(sizeof(int64_t) * (size_t)B * 2 * N);
    int64_t *o2 = malloc(sizeof(int64_t) * (size_t)B * 2 * N);
    for (size_t i = 0; i < key_n; i++)
        bsk[i] = rand_class(log_q);
    for (size_t i = 0; i < (size_t)B * n_iter * m; i++)
        idx[i] = (int32_t)(next_u64() % (uint64_t)(2 * N));
    for (int i = 0; i < 2 * N; i++)
        lut[i] = rand_class(log_q);
    for (int b = 0; b < B; b++)
        memcpy(lut_rep + (size_t)b * 2 * N, lut, sizeof(int64_t) * 2 * N);
    tfhe_exact_blind_rotate(lut, 0, bsk, idx, o1, B, n_iter, m, N, L, log_beta, log_q);
    tfhe_exact_blind_rotate(lut_rep, 2 * N, bsk, idx, o2, B, n_iter, m, N, L, log_beta, log_q);
    int bad = memcmp(o1, o2, sizeof(int64_t) * (size_t)B * 2 * N) != 0;
    tfhe_exact_blind_rotate(lut, 0, zero, idx, o1, B, n_iter, m, N, L, log_beta, log_q);
    for (size_t i = 0; i < (size_t)B * 2 * N; i++)
        bad += o1[i] != 0;
    free(bsk);
    free(zero);
    free(idx);
    free(lut);
    free(lut_rep);
    free(o1);
    free(o2);
    if (bad)
        printf("blind_rotate B=%d N=%d L=%d log_q=%d: mismatch\n", B, N, L, log_q);
    return bad;
}

int main(void)
{
    int bad = 0;
    bad += check_mul(64, 32);
    bad += check_mul(128, 64);
    bad += check_mul(256, 20);
    bad += check_blind_rotate(3, 4, 8, 64, 3, 7, 32);
    bad += check_blind_rotate(2, 3, 4, 128, 2, 13, 64);
    bad += check_blind_rotate(1, 2, 2, 32, 4, 4, 20);
    printf(bad ? "sanitize check FAILED\n" : "sanitize check ok\n");
    return bad ? 1 : 0;
}
