"""Pins of the LSTM-path oracle (oracle/lstm_oracle.py) against the
mathematics: closed forms, exact identities with noise-free keys and the
int128 schoolbook product.  Each check is chosen so that a plausible slip (a
dropped sign, a wrong automorphism exponent, a transposed operand, an
off-by-one digit level) fails it."""
import numpy as np
import pytest

from oracle import exact
from oracle import lstm_oracle as L
from oracle import tfhe_oracle as O


@pytest.mark.parametrize("log_q", [20, 32, 48])
def test_digits_equal_float_rule_where_float_is_exact(log_q):
    rng = np.random.default_rng(log_q)
    x = L.uniform(rng, 4000, log_q)
    for log_beta, l in [(2, log_q // 2), (7, 2), (4, log_q // 4)]:
        want = O.decompose_float(x.astype(np.float64), float(2 ** log_beta), l, float(2 ** log_q))
        assert np.array_equal(L.digits(x, log_beta, l, log_q), want.astype(np.int64))


@pytest.mark.parametrize("log_beta,l", [(2, 20), (8, 4), (2, 32), (16, 4), (13, 2)])
def test_digits_recompose_at_q64(log_beta, l):
    rng = np.random.default_rng(l)
    x = L.uniform(rng, 20000, 64)
    d = L.digits(x, log_beta, l, 64)
    assert np.abs(d).max() <= 2 ** (log_beta - 1)
    err = L.tmod(x - L.recompose(d, log_beta, 64), 64)
    bound = 2 ** (64 - l * log_beta - 1) if l * log_beta < 64 else 0
    assert np.abs(err).max() <= bound
    if l * log_beta == 64:
        assert np.array_equal(L.recompose(d, log_beta, 64), x)


def _residual_at(x, S, log_q):
    """The closed form csrc/torus_device.h's residual_at computes: the centred
    remainder of x mod 2^S, the tie 2^(S-1) negative iff bit S of x is set."""
    if S >= log_q:
        return x.copy()
    if S <= 0:
        return np.zeros_like(x)
    u = x.astype(np.uint64)
    rem = u & np.uint64((1 << S) - 1)
    half = np.uint64(1 << (S - 1))
    up = (rem > half) | ((rem == half) & (((u >> np.uint64(S)) & np.uint64(1)) == 1))
    return np.where(up, (rem - np.uint64(1 << S)).view(np.int64), rem.view(np.int64))


@pytest.mark.parametrize("log_q,log_beta,l", [(64, 2, 20), (64, 13, 2), (64, 8, 8), (32, 7, 3),
                                              (48, 4, 12), (20, 3, 7), (24, 5, 6)])
def test_digit_recursion_restarts_from_closed_form_residual(log_q, log_beta, l):
    """The trace-packing kernel starts the digit recursion of each gadget-level
    pair at level k from the closed-form residual; continuing the recursion
    from it must give the oracle's digits of levels k, k+1, ..., also where
    round-half ties are planted at every level (a residual that ignored the
    tie bit, or kept one bit too few, fails here)."""
    rng = np.random.default_rng(log_q * 100 + log_beta)
    x = L.uniform(rng, 4000, log_q).astype(np.uint64)
    for j in range(l):                        # ties at level j in a slice of the inputs
        s = log_q - (j + 1) * log_beta
        if s >= 1:
            sl = slice(j * 150, j * 150 + 300)
            x[sl] = (x[sl] & ~np.uint64((1 << s) - 1)) | np.uint64(1 << (s - 1))
    x = L.tmod(x.view(np.int64), log_q)
    D = L.digits(x, log_beta, l, log_q)
    with np.errstate(over="ignore"):
        for k in range(l):
            r = _residual_at(x, log_q - k * log_beta, log_q)
            for j in range(k, l):
                s = log_q - (j + 1) * log_beta
                if s <= 0:
                    d, r = r, np.zeros_like(r)
                else:
                    d = L._rhe_shift(r, s)
                    r = (r.astype(np.uint64) - (d.astype(np.uint64) << np.uint64(s))).view(np.int64)
                assert np.array_equal(d, D[j]), (k, j)


def _carries(x, log_q):
    """csrc/torus_device.h's carries_of: bit p = up(p + 1), the round-half-even
    carry of the residual at weight 2^(p+1): bit p of x and (bit p+1, or p
    above the lowest set bit), the sign bit or-ed in at p = log_q - 1."""
    u = x.astype(np.uint64)
    low = u & (np.uint64(0) - u)
    above = ~(low | (low - np.uint64(1)))
    return (u & ((u >> np.uint64(1)) | above)) | (u & np.uint64(1 << (log_q - 1)))


def _digit_from_bits(x, j, log_beta, log_q):
    """csrc/torus_device.h's digit_level / digit_at_level for level j of the
    float-rule recursion, on 32-bit words as the device has them: with S =
    log_q - j*log_beta, S1 = max(S - log_beta, 0) and w = S - S1, uu = x >> S1
    holds the field in bits [0, w) and cc = cv >> (S1 - 1) (cv << 1 at S1 = 0)
    holds up(S1) in bit 0 and up(S) in bit w; the field inserted into cc and
    read as a signed (w+1)-bit number is field - up(S) * 2^w, and bit 0 of cc
    adds up(S1) (bit p of the carry word cv is up(p + 1))."""
    u = x.astype(np.uint64)
    cv = _carries(x, log_q)
    S = log_q - j * log_beta
    if S <= 0:
        return np.zeros(u.shape, np.int64)
    S1 = max(S - log_beta, 0)
    w = S - S1
    m32 = np.uint64(0xFFFFFFFF)
    uu = (u >> np.uint64(S1)) & m32
    cc = ((cv >> np.uint64(S1 - 1)) if S1 > 0 else (cv << np.uint64(1))) & m32
    fm = np.uint64((1 << w) - 1)
    ins = ((uu & fm) | (cc & ~fm)) & np.uint64((1 << (w + 1)) - 1)     # bits [0, w]
    signed = ins.astype(np.int64) - ((ins >> np.uint64(w)).astype(np.int64) << (w + 1))
    return signed + (cc & np.uint64(1)).astype(np.int64)


@pytest.mark.parametrize("log_q,log_beta,l", [(64, 2, 20), (64, 13, 2), (64, 8, 8), (32, 7, 3),
                                              (48, 4, 12), (20, 3, 7), (24, 5, 6), (64, 2, 32),
                                              (32, 4, 8), (16, 1, 16), (12, 5, 4)])
def test_digits_from_bit_fields_equal_recursion(log_q, log_beta, l):
    """The trace-packing kernel reads each digit of the float rule straight
    from bit fields of the centred class and its carry word (no recursion): equal
    to the oracle's digits at every level, with round-half ties planted at
    every level and the all-zero / sign-boundary classes included.  A carry
    that ignored the tie bit, the sticky lower bits, or the sign at the top
    level fails here, and so does a carry word that misses the sign at
    log_q = 64 (the class -2^63 is planted)."""
    rng = np.random.default_rng(log_q * 1000 + log_beta * 10 + l)
    x = L.uniform(rng, 6000, log_q).astype(np.uint64)
    for j in range(l):
        s = log_q - (j + 1) * log_beta
        if s >= 1:
            sl = slice(j * 150, j * 150 + 300)
            x[sl] = (x[sl] & ~np.uint64((1 << s) - 1)) | np.uint64(1 << (s - 1))
            sl2 = slice(j * 150 + 3000, j * 150 + 3050)    # ties with the tie bit's neighbour set
            x[sl2] = (x[sl2] & ~np.uint64((1 << s) - 1)) | np.uint64(1 << (s - 1)) | np.uint64(1 << s)
    x = L.tmod(x.view(np.int64), log_q)
    x[:4] = [0, -(1 << (log_q - 1)) if log_q < 64 else np.iinfo(np.int64).min, -1, 1]
    D = L.digits(x, log_beta, l, log_q)
    for j in range(l):
        assert np.array_equal(_digit_from_bits(x, j, log_beta, log_q), D[j]), j


def test_tmod_and_rhe():
    assert L.tmod(np.array([2 ** 31, -2 ** 31 - 1, 5]), 32).tolist() == [-2 ** 31, 2 ** 31 - 1, 5]
    r = np.array([5, 6, 7, -5, -6, -7, 12, -12], np.int64)       # /4: 1.25 1.5 1.75 ...
    assert L._rhe_shift(r, 2).tolist() == [1, 2, 2, -1, -2, -2, 3, -3]
    assert L._rhe_shift(np.array([2, 6, -2, -6]), 2).tolist() == [0, 2, 0, -2]   # ties to even


@pytest.mark.parametrize("N,log_q", [(64, 64), (256, 64), (256, 32), (1024, 64)])
def test_negacyclic_mul_exact_vs_int128_schoolbook(N, log_q):
    rng = np.random.default_rng(N)
    a = L.uniform(rng, N, log_q)
    d = rng.integers(-2 ** 12, 2 ** 12, N)
    assert np.array_equal(L.negacyclic_mul_exact(a, d, log_q), exact.negacyclic_mul(a, d, log_q))


def test_apply_automorphism_is_a_ring_homomorphism():
    rng = np.random.default_rng(3)
    N = 128
    a = rng.integers(-50, 50, N)
    b = rng.integers(-50, 50, N)
    ab = exact.negacyclic_mul(a, b, 62)
    for k in (3, 5, 2 * N - 1, N + 1, 65):
        lhs = L.apply_automorphism(ab, k)
        rhs = exact.negacyclic_mul(L.apply_automorphism(a, k), L.apply_automorphism(b, k), 62)
        assert np.array_equal(lhs, rhs), k
    assert np.array_equal(L.apply_automorphism(L.apply_automorphism(a, 5), 3),
                          L.apply_automorphism(a, 15))


def test_apply_automorphism_monomial_images():
    N = 16
    for j in range(N):
        for k in (3, 5, 17, 31):
            x = np.zeros(N, np.int64)
            x[j] = 1
            y = L.apply_automorphism(x, k)
            e = (j * k) % (2 * N)
            want = np.zeros(N, np.int64)
            want[e % N] = 1 if e < N else -1
            assert np.array_equal(y, want)


def test_weights_to_polynomial_gives_dot_product():
    rng = np.random.default_rng(5)
    N, K, out = 64, 20, 7
    W = rng.integers(-7, 8, (K, out))
    P = L.weights_to_polynomial(W, N)
    m = np.zeros(N, np.int64)
    m[:K] = rng.integers(-100, 100, K)
    for o in range(out):
        prod = exact.negacyclic_mul(P[o], m, 62)
        assert prod[0] == int(W[:, o] @ m[:K])


@pytest.mark.parametrize("log_q", [32, 64])
def test_linear_extract_phase_identity(log_q):
    """phase(Linear(ct))_o == (W_o * phase(ct))_0 + bias_o, exactly."""
    rng = np.random.default_rng(log_q)
    N, K, out = 128, 40, 6
    S = rng.integers(0, 2, N)
    m = np.zeros(N, np.int64)
    m[:K] = rng.integers(-1000, 1000, K) << (log_q - 20)
    A, b = L.rlwe_encrypt(rng, m, S, log_q, 2.0 ** 4)
    W = rng.integers(-7, 8, (K, out))
    P = L.weights_to_polynomial(W, N)
    bias = rng.integers(-2 ** 10, 2 ** 10, out)
    masks, bodies = L.linear_extract((A, b), P, bias, log_q)
    ph = L.lwe_phase(masks, bodies, S, log_q)
    ph_ct = L.rlwe_phase(A, b, S, log_q)
    for o in range(out):
        want = L.tmod(L.negacyclic_mul_exact(ph_ct, P[o], log_q)[0] + bias[o], log_q)
        assert ph[o] == want


def test_trace_pack_noise_free_is_exact():
    """Zero-noise Galois keys and a full decomposition: the packed RLWE's
    phase is N * phase(modulus-switched LWE) on coefficient 0, zero elsewhere."""
    rng = np.random.default_rng(11)
    N, log_q, lb, l = 64, 32, 4, 8
    S = rng.integers(0, 2, N)
    gk = L.galois_key(rng, S, log_q, lb, l, 0.0)
    for _ in range(3):
        a, b = L.lwe_encrypt(rng, int(rng.integers(-2 ** 20, 2 ** 20)) << 8, S, log_q, 0.0)
        A, Bp = L.trace_pack(a, int(b), gk, lb, l, log_q)
        ph = L.rlwe_phase(A, Bp, S, log_q)
        a2 = L.modulus_switch(a, 6)
        b2 = L.modulus_switch(np.array([b]), 6)[0]
        want = L.tmod(N * L.lwe_phase(a2, b2, S, log_q), log_q)
        assert ph[0] == want
        assert np.all(ph[1:] == 0)


def test_trace_pack_noisy_keys_bounded():
    rng = np.random.default_rng(12)
    N, log_q, lb, l = 64, 64, 2, 20
    S = rng.integers(0, 2, N)
    gk = L.galois_key(rng, S, log_q, lb, l, 2.0 ** 17)
    mu = 3 << 56
    a, b = L.lwe_encrypt(rng, mu, S, log_q, 2.0 ** 17)
    A, Bp = L.trace_pack(a, int(b), gk, lb, l, log_q)
    ph = L.rlwe_phase(A, Bp, S, log_q)
    # the modulus switch divides by N and the trace multiplies by N: mu again
    assert abs(int(L.tmod(ph[0] - np.int64(mu), 64))) < 2 ** 50
    assert np.abs(ph[1:]).max() < 2 ** 50


def test_packing_noise_free_is_exact():
    rng = np.random.default_rng(13)
    N, n, log_q, lb, l = 64, 48, 32, 8, 4
    S = rng.integers(0, 2, N)
    s = rng.integers(0, 2, n)
    ksk = L.packing_ksk(rng, s, S, log_q, lb, l, 0.0)
    n_lwes = 20
    mu = rng.integers(-2 ** 30, 2 ** 30, n_lwes)
    a, b = L.lwe_encrypt(rng, mu, s, log_q, 2.0 ** 3)
    A, Bp = L.packing(a, b, ksk, lb, l, log_q)
    ph = L.rlwe_phase(A, Bp, S, log_q)
    assert np.array_equal(ph[:n_lwes], L.lwe_phase(a, b, s, log_q))
    assert np.all(ph[n_lwes:] == 0)


def test_lut_product_phase_identity():
    rng = np.random.default_rng(14)
    N, log_q, G, nl = 64, 64, 3, 4
    S = rng.integers(0, 2, N)
    m = rng.integers(-2 ** 40, 2 ** 40, (nl * G, N))
    A, b = L.rlwe_encrypt(rng, m, S, log_q, 2.0 ** 17)
    lut = rng.integers(-1, 2, (nl * G, N))
    out = L.lut_product_sum(np.stack([A, b], axis=1), lut, G, nl, log_q)
    for g in range(G):
        ph = L.rlwe_phase(out[g, 0], out[g, 1], S, log_q)
        want = sum(L.negacyclic_mul_exact(L.rlwe_phase(A[k * G + g], b[k * G + g], S, log_q),
                                          lut[k * G + g], log_q).astype(np.uint64)
                   for k in range(nl))
        assert np.array_equal(ph, L.tmod(np.asarray(want).astype(np.int64), log_q))


@pytest.mark.parametrize("log_q,lb,l", [(64, 2, 32), (32, 4, 8)])
def test_key_switch_int_noise_free_keeps_phase(log_q, lb, l):
    rng = np.random.default_rng(15)
    n_in, n_out = 96, 40
    s_old, s_new = rng.integers(0, 2, n_in), rng.integers(0, 2, n_out)
    ka, kb = L.keyswitch_key(rng, s_new, s_old, log_q, lb, l, 0.0)
    a, b = L.lwe_encrypt(rng, rng.integers(-2 ** 20, 2 ** 20, 7) << (log_q - 24), s_old, log_q, 4.0)
    a2, b2 = L.key_switch_int(ka, kb, a, b, log_q, lb, l)
    assert np.array_equal(L.lwe_phase(a2, b2, s_new, log_q), L.lwe_phase(a, b, s_old, log_q))


def test_rotate_extract_int_matches_tfhe_oracle():
    rng = np.random.default_rng(16)
    N = 64
    br = L.uniform(rng, (3, 2, N), 32)
    body = L.uniform(rng, 3, 32)
    masks, bodies = L.rotate_extract_int(br, body, 32, 4, 4)
    for i in range(3):
        bms = int(O.body_rotation(float(body[i]), 32, N, 4))
        A, Bp = O.rotate_ciphertext((br[i, 0], br[i, 1]), -bms)
        for k in range(4):
            a, bb = O.sample_extract((A, Bp), k)
            assert np.array_equal(masks[i, k], O.centered_mod(a, 32))
            assert bodies[i, k] == O.centered_mod(bb, 32)[0]


def test_rotation_indices_int_matches_float_rule_at_q32():
    rng = np.random.default_rng(17)
    pk = L.uniform(rng, (5, 30), 32)
    assert np.array_equal(L.rotation_indices_int(pk, 32, 256, 3, 4),
                          O.rotation_indices(pk, 32, 256, 3, 4))
