"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors or known-answer tests for this path
(its only check is a printed mean-error test, test_collapse.py:186,229), so
each oracle function is pinned by closed forms, invariants, brute force on
small inputs and the exact int128 C oracle.  Planted mistakes are checked to
fail the same pins.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import exact as X
from oracle import tfhe_oracle as O
from tests.tfhe_cases import decrypt_outputs, expected_mod, make_case


def test_all_binary_vectors_closed_form():
    for c in range(1, 6):
        v = O.all_binary_vectors(c)
        grids = np.meshgrid(*([np.array([0, 1])] * c), indexing="ij")
        np.testing.assert_array_equal(v, np.stack(grids, -1).reshape(-1, c))
        # row p is p in binary, MSB first
        assert all(int("".join(map(str, row)), 2) == p for p, row in enumerate(v))


def test_centered_mod_brute_force():
    x = np.arange(-600, 600)
    r = O.centered_mod(x, 8)
    assert np.all((r >= -128) & (r < 128))
    assert np.all(np.mod(r - x, 256) == 0)


@pytest.mark.parametrize("log_q,log_beta,L", [(8, 2, 3), (8, 3, 2), (10, 2, 5), (9, 3, 3), (8, 4, 2)])
def test_decompose_int_exhaustive(log_q, log_beta, L):
    """dec_fft.cuh:4-31: digits in [-beta/2, beta/2) recompose to x rounded
    (half up) to a multiple of q/beta^L, mod q."""
    q = 1 << log_q
    x = np.arange(-q // 2, q // 2)
    d = O.decompose_int(x, log_beta, log_q, L)
    beta = 1 << log_beta
    assert d.min() >= -beta // 2 and d.max() < beta // 2
    rec = sum(d[i] * (1 << (log_q - (i + 1) * log_beta)) for i in range(L))
    unit = 1 << (log_q - L * log_beta)
    want = ((x + unit // 2) // unit) * unit if unit > 1 else x
    assert np.all(np.mod(rec - want, q) == 0)


def test_decompose_int_planted_mistake_fails():
    """Dropping the rounding term of the offset breaks the invariant."""
    log_q, log_beta, L = 8, 2, 3
    x = np.arange(-128, 128)
    off = sum(1 << (log_q - i * log_beta + log_beta - 1) for i in range(1, L + 1))
    st = x.astype(np.uint64) + np.uint64(off)             # no rounding term
    d = [((st >> np.uint64(log_q - i * log_beta)) & np.uint64(3)).astype(np.int64) - 2
         for i in range(1, L + 1)]
    rec = sum(d[i] * (1 << (log_q - (i + 1) * log_beta)) for i in range(L))
    want = ((x + 1) // 2) * 2
    assert not np.all(np.mod(rec - want, 256) == 0)


def test_decompose_float_matches_int_off_ties():
    """polynomial_jax.py:90-111 recomposes to round-half-even; it agrees with
    the integer rule everywhere except on exact ties."""
    log_q, log_beta, L = 12, 3, 3
    q = 1 << log_q
    x = np.arange(-q // 2, q // 2)
    df = O.decompose_float(x.astype(np.float64), 2.0 ** log_beta, L, float(q)).astype(np.int64)
    assert np.abs(df).max() <= (1 << (log_beta - 1))
    rec = sum(df[i] * (1 << (log_q - (i + 1) * log_beta)) for i in range(L))
    unit = 1 << (log_q - L * log_beta)
    assert np.all(np.abs(O.centered_mod(rec - x, log_q)) <= unit // 2)
    di = O.decompose_int(x, log_beta, log_q, L)
    reci = sum(di[i] * (1 << (log_q - (i + 1) * log_beta)) for i in range(L))
    same = np.mod(rec - reci, q) == 0
    ties = np.mod(x, unit) == unit // 2
    assert np.all(same | ties)


@pytest.mark.parametrize("N", [8, 16, 64])
def test_negacyclic_fft_closed_form(N):
    rng = np.random.default_rng(N)
    p = rng.integers(-1000, 1000, N).astype(np.float64)
    F = O.negacyclic_fft(p)
    zeta = np.exp(-1j * np.pi / N)
    direct = np.array([sum(p[k] * zeta ** (k * (2 * j + 1)) for k in range(N)) for j in range(N // 2)])
    np.testing.assert_allclose(F, direct, rtol=1e-10, atol=1e-6)
    np.testing.assert_array_equal(O.negacyclic_ifft(F), p)


def test_negacyclic_fft_product_is_exact_product():
    rng = np.random.default_rng(1)
    N = 64
    a = rng.integers(-2 ** 20, 2 ** 20, N)
    b = rng.integers(-64, 64, N)
    prod = O.negacyclic_ifft(O.negacyclic_fft(a) * O.negacyclic_fft(b)).astype(np.int64)
    np.testing.assert_array_equal(prod, O.negacyclic_mul_direct(a, b).astype(np.int64))


def test_exact_c_oracle_matches_python_ints():
    rng = np.random.default_rng(2)
    for N in (16, 128):
        a = rng.integers(-2 ** 31, 2 ** 31, N)
        b = rng.integers(-2 ** 31, 2 ** 31, N)
        np.testing.assert_array_equal(X.negacyclic_mul(a, b, 32), O.negacyclic_mul_direct(a, b, 32))


def test_negacyclic_mul_binary_matches_direct():
    rng = np.random.default_rng(3)
    a = rng.integers(-2 ** 31, 2 ** 31, (3, 64))
    s = rng.integers(0, 2, 64)
    ex = np.array([O.negacyclic_mul_direct(r, s) for r in a]).astype(np.int64)
    np.testing.assert_array_equal(O.negacyclic_mul_binary(a, s), ex)


def test_coef_rotation_power_series():
    N = 16
    p = np.arange(1, N + 1)
    for alpha in range(-2 * N, 2 * N + 1, 3):
        mono = np.zeros(N, np.int64)
        e = alpha % (2 * N)
        mono[e % N] = 1 if e < N else -1
        np.testing.assert_array_equal(O.coef_rotation(p, alpha),
                                      O.negacyclic_mul_direct(p, mono).astype(np.int64))
    # X^a X^b = X^(a+b)
    np.testing.assert_array_equal(O.coef_rotation(O.coef_rotation(p, 5), -11), O.coef_rotation(p, -6))


def test_rotation_indices_brute_force_with_ties():
    rng = np.random.default_rng(4)
    N, n, c, log_q = 64, 12, 3, 16
    q = 1 << log_q
    for n_lut in (1, 2, 4):
        pk = rng.integers(-q // 2, q // 2, (20, n))
        unit = q * n_lut // (4 * N)          # odd multiples sit on x.5
        pk[:5] = (2 * rng.integers(-20, 20, (5, n)) + 1) * unit
        idx = O.rotation_indices(pk.astype(np.float64), log_q, N, c, n_lut)
        kron = O.all_binary_vectors(c)
        for b in range(20):
            for g in range(n // c):
                for p in range(2 ** c):
                    val = sum(int(pk[b, g * c + j]) * int(kron[p, j]) for j in range(c))
                    t = Fraction(2 * N * val, q * n_lut)
                    r = round(t)                      # Python round = half even
                    assert idx[b, g, p] == (r * n_lut) % (2 * N)


def test_sample_extract_identity():
    rng = np.random.default_rng(5)
    N = 32
    S = rng.integers(0, 2, N)
    A = rng.integers(-2 ** 31, 2 ** 31, N)
    B = rng.integers(-2 ** 31, 2 ** 31, N)
    phase = O.centered_mod(B - O.negacyclic_mul_binary(A, S), 32)
    for k in range(N):
        a, b = O.sample_extract((A, B), k)
        assert O.centered_mod(b[0] - a @ S, 32) == phase[k]


def test_blind_rotate_f64_matches_exact_and_luts():
    case = make_case(N=128, n=24, L=3, log_beta=7, B=16, msg_bits=3, seed=0, sigma_bsk=2.0 ** 6)
    idx = O.rotation_indices(case.a, 32, case.N, 3, 1)
    br_x = X.blind_rotate(case.lut, case.bsk_coef_k, idx, 7, 32, 3)
    br_f = O.blind_rotate_f64(np.broadcast_to(case.lut, (16, 2, case.N)), case.bsk_fft_k, idx, 7, 32, 3)
    assert np.abs(O.centered_mod(br_f - br_x, 32)).max() <= 4
    masks, bodies = O.cuboot_merge_tail(br_x, case.b, 32, 1)
    dec, _ = decrypt_outputs(case, masks, bodies)
    np.testing.assert_array_equal(dec, expected_mod(case))


def test_blind_rotate_noise_free_kat():
    """Noise-free keys and input: the PBS output is exactly LUT(m) on the
    constant coefficient, up to the decomposition rounding."""
    case = make_case(N=64, n=12, L=4, log_beta=8, B=24, msg_bits=3, seed=3, sigma_bsk=0.0,
                     sigma_lwe=0.0)
    idx = O.rotation_indices(case.a, 32, case.N, 3, 1)
    br = X.blind_rotate(case.lut, case.bsk_coef_k, idx, 8, 32, 4)
    masks, bodies = O.cuboot_merge_tail(br, case.b, 32, 1)
    _, err = decrypt_outputs(case, masks, bodies)
    assert np.all(err == 0)


def test_many_lut_outputs():
    case = make_case(N=256, n=18, L=3, log_beta=7, B=32, n_lut=4, msg_bits=3, seed=6,
                     sigma_bsk=2.0 ** 4, sigma_lwe=0.0, kind="many")
    idx = O.rotation_indices(case.a, 32, case.N, 3, 4)
    br = X.blind_rotate(case.lut, case.bsk_coef_k, idx, 7, 32, 3)
    masks, bodies = O.cuboot_merge_tail(br, case.b, 32, 4)
    dec, _ = decrypt_outputs(case, masks, bodies)
    np.testing.assert_array_equal(dec, expected_mod(case))


def test_key_switch_preserves_phase():
    rng = np.random.default_rng(7)
    n_in, n_out, lb, l = 128, 40, 4, 5
    s_old = rng.integers(0, 2, n_in)
    s_new = rng.integers(0, 2, n_out)
    ka, kb = O.keyswitch_key(rng, s_new, s_old, 32, lb, l, 0.0)
    a = rng.integers(-2 ** 31, 2 ** 31, (10, n_in))
    b = rng.integers(-2 ** 31, 2 ** 31, 10)
    na, nb = O.key_switch_lwe(ka, kb, a, b, 32, lb, l)
    ph_old = O.centered_mod(b - a @ s_old, 32)
    ph_new = O.centered_mod(nb - na @ s_new, 32)
    # noise-free key: only the dropped low digits remain, |err| <= n_in*q/beta^l/2
    assert np.abs(O.centered_mod(ph_new - ph_old, 32)).max() <= n_in * 2 ** (32 - lb * l - 1)


def test_get_plain_LUT_closed_form():
    N, n_lut, s, gamma = 64, 4, 2.0, 0.3
    for idx in range(n_lut):
        T = O.get_plain_LUT(N, n_lut, idx, s, gamma)
        sub = N // n_lut
        vals = np.where(np.linspace(-s, s, sub, endpoint=False) >= gamma, 1.0, 0.0)
        # signed phase psi in [-N/2, N/2) reads complete[psi + N/2]
        for psi in range(-N // 2, N // 2):
            e = psi % (2 * N)
            coef0 = O.coef_rotation(T, -e)[0]
            full = np.zeros(N)
            full[np.arange(sub) * n_lut + idx] = vals
            assert coef0 == full[psi + N // 2]


@pytest.mark.parametrize("kind,n_lut,msg_bits,N,n", [("identity", 1, 3, 256, 24), ("many", 4, 4, 512, 36),
                                                   ("many", 2, 3, 128, 12)])
def test_noise_free_pbs_is_the_lut_at_the_switched_phase(kind, n_lut, msg_bits, N, n):
    """The exact (int128) oracle's PBS decrypts to the test polynomial read at
    the modulus-switched phase, sum_g idx[g, p_g] - b_ms (tests/tfhe_cases.py):
    the expectation the GPU tests use at full size.  Inputs with large LWE
    noise make many outputs miss LUT(m), so the check is not the trivial one."""
    from oracle import exact as X
    from tests.tfhe_cases import decrypt_outputs, expected_mod, lut_at_switched_phase, make_case
    case = make_case(N=N, n=n, L=3, log_beta=7, B=40, n_lut=n_lut, msg_bits=msg_bits, seed=77,
                     sigma_bsk=2.0 ** 4, sigma_lwe=2.0 ** 27, kind=kind)
    idx = O.rotation_indices(case.a, 32, N, 3, n_lut)
    br = X.blind_rotate(case.lut, case.bsk_coef_k, idx, 7, 32, 3)
    em, eb = O.cuboot_merge_tail(br, case.b, 32, n_lut)
    dec, _ = decrypt_outputs(case, em, eb)
    want = lut_at_switched_phase(case)
    np.testing.assert_array_equal(dec, want)
    assert (want != expected_mod(case)).any()        # the noisy inputs do miss LUT(m)


def test_c_oracle_clean_under_address_and_ub_sanitizers():
    """The exact C oracle built with -fsanitize=address,undefined (host code,
    gcc) runs its own checks (products against a schoolbook, broadcast vs
    per-ciphertext blind rotation, zero key) without a sanitizer report."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc/make not available")
    root = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))
    r = subprocess.run(["make", "-s", "-f", "oracle/sanitize.mk"], cwd=root, capture_output=True,
                       text=True, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "sanitize check ok" in out
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out
