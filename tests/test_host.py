"""Host-side pieces of the package (key generation, encryption, layouts, LUT
builders) against the oracle."""
import numpy as np
import torch

from oracle import tfhe_oracle as O
from tests.tfhe_cases import MANY_LUT_FNS


def test_negacyclic_mul_small_exact():
    from schemas.polynomial_jax import negacyclic_mul_small
    rng = np.random.default_rng(0)
    for N, log_q in [(64, 32), (256, 32), (128, 64), (64, 20)]:
        a = rng.integers(-2 ** 31, 2 ** 31, (4, N))
        s = rng.integers(-1, 2, N)
        want = np.array([O.negacyclic_mul_direct(r, s, log_q if log_q < 64 else None) for r in a])
        if log_q == 64:
            want = np.array([[int(v) for v in row] for row in want], dtype=object)
            want = np.array([[((v + 2 ** 63) % 2 ** 64) - 2 ** 63 for v in row] for row in want],
                            dtype=np.int64)
        np.testing.assert_array_equal(negacyclic_mul_small(a, s, log_q), want)


def test_rlwe_encrypt_decrypt():
    from schemas import RLWE_jax as R
    params = dict(q=2.0 ** 32, t=16, degree=128, sigma=2.0 ** 10)
    sk = R.sample_sk(params, 1)
    m = np.random.default_rng(2).integers(-8, 8, (3, 128)).astype(np.float64)
    A, b = R.encrypt(m * 2.0 ** 28, sk, params, 3)
    assert np.all(np.mod(R.decrypt((A, b), sk, params) - m, 16) == 0)
    ph = O.rlwe_phase(A.astype(np.int64), b.astype(np.int64), sk.astype(np.int64), 32)
    assert np.abs(O.centered_mod(ph - (m * 2 ** 28).astype(np.int64), 32)).max() < 2 ** 10 * 8


def test_lwe_encrypt_decrypt_and_ksk():
    from schemas import LWE_jax as LW
    rng = np.random.default_rng(4)
    params = dict(q=2.0 ** 32, t=16, degree=64, sigma=2.0 ** 8, beta_ks=16, l_ks=5)
    s_new = rng.integers(0, 2, 40)
    s_old = rng.integers(0, 2, 64)
    m = rng.integers(-8, 8, 10).astype(np.float64)
    a, b = LW.encrypt_LWE(m * 2.0 ** 28, s_new, params, rng)
    assert np.all(np.mod(LW.decrypt_LWE((a, b), s_new, params) - m, 16) == 0)
    ka, kb = LW.get_keyswitch_sk_LWE(s_new, s_old, params, rng)
    assert ka.shape == (5, 64, 40) and kb.shape == (5, 64, 1)
    # KSK row (j, i) decrypts to -s_old[i] * q / beta^(j+1)
    ph = O.centered_mod(kb[..., 0].astype(np.int64) - ka.astype(np.int64) @ s_new, 32)
    want = O.centered_mod(-s_old[None, :] * (2 ** 32 // 16 ** np.arange(1, 6))[:, None], 32)
    assert np.abs(O.centered_mod(ph - want, 32)).max() <= 2 ** 8 * 8


def test_bsk_keygen_layout_and_phases():
    from cuTFHE.multiply import prepare_bsk
    from schemas import RLWE_jax as R
    rng = np.random.default_rng(5)
    N, n, c, L, lb = 64, 12, 3, 3, 7
    params = dict(q=2.0 ** 32, beta_bs=2 ** lb, l_bs=L, sigma=0.0, degree=N)
    S = rng.integers(0, 2, N).astype(np.float64)
    s = rng.integers(0, 2, n)
    BSK = R.get_boostrapping_key_collapse(S, s, params, rng, c)
    coef = np.asarray([[np.asarray(x) for x in side] for side in BSK]).astype(np.int64)
    assert coef.shape == (2, 2, n // c * 2 ** c, L, N)
    kron = O.all_binary_vectors(c)
    ind = np.all(s.reshape(-1, c)[:, None, :] == kron[None], -1).astype(np.int64).ravel()
    for i in range(L):
        g = 2 ** (32 - (i + 1) * lb)
        ph_left = O.rlwe_phase(coef[0, 0, :, i], coef[0, 1, :, i], S.astype(np.int64), 32)
        np.testing.assert_array_equal(ph_left, O.centered_mod(-np.outer(ind, S.astype(np.int64)) * g, 32))
        ph_right = O.rlwe_phase(coef[1, 0, :, i], coef[1, 1, :, i], S.astype(np.int64), 32)
        want = np.zeros_like(ph_right)
        want[:, 0] = ind * g
        np.testing.assert_array_equal(ph_right, O.centered_mod(want, 32))
    # Fourier layout and kernel reshape equal the oracle's
    ordered = R.bsk_to_fourier(BSK)
    np.testing.assert_allclose(ordered, O.bsk_to_fourier_ordered(coef), rtol=0, atol=1e-3)
    k = prepare_bsk(torch.as_tensor(ordered), n, c).numpy()
    np.testing.assert_array_equal(k, O.reshape_bsk_for_kernel(ordered, n // c, c))


def test_lut_builders_match_oracle():
    from nn_tfhe.utils_spike import get_plain_LUT, many_lut_polynomial
    for idx in range(4):
        np.testing.assert_array_equal(get_plain_LUT(256, 4, idx, 1.5, -0.2),
                                      O.get_plain_LUT(256, 4, idx, 1.5, -0.2))
    for bits in (3, 4):
        np.testing.assert_array_equal(many_lut_polynomial(2048, MANY_LUT_FNS, bits, 2.0 ** 27),
                                      O.lut_from_function(2048, 4, MANY_LUT_FNS, bits, 2 ** 27))


def test_all_binary_vectors_product_matches_oracle():
    from cuTFHE.utils import all_binary_vectors
    for c in range(1, 5):
        np.testing.assert_array_equal(all_binary_vectors(c), O.all_binary_vectors(c))


def test_tuning_table_round_trip(tmp_path):
    from cuTFHE import multiply
    saved = dict(multiply._TUNED)
    try:
        multiply._TUNED.clear()
        multiply._TUNED[(2048, 2, 234, 8, 6, False, ("plain", "batched", "rows"))] = "rows"
        multiply._TUNED[(2048, 3, 250, 8, 14, True, ("plain", "rows"))] = "plain"
        path = str(tmp_path / "tune.json")
        multiply.save_tuning(path)
        want = dict(multiply._TUNED)
        multiply._TUNED.clear()
        assert multiply.load_tuning(path) == 2
        assert multiply._TUNED == want
    finally:
        multiply._TUNED.clear()
        multiply._TUNED.update(saved)


def test_plaintext_step_lut_negacyclic_extension():
    """The plaintext gate equals the test polynomial get_plain_LUT builds,
    read the way the blind rotation reads it (coefficient 0 of X^-phi * T,
    negacyclic), over the whole torus and not only [-s, s)."""
    from nn_tfhe.plaintext import lut_margin, step_lut
    from nn_tfhe.utils_spike import get_plain_LUT
    N, n_lut, s, gamma = 256, 4, 4.0, 0.7
    for k in range(n_lut):
        T = np.asarray(get_plain_LUT(N, n_lut, k, s, gamma), np.float64)
        for j in range(2 * N // n_lut):                # LUT entries over the whole torus
            phi = j * n_lut                            # rounded phase, multiple of n_lut
            read = T[phi + k] if phi + k < N else -T[phi + k - N]
            pre = j * 2 * s / (N // n_lut)             # phase j entries above zero
            if lut_margin(pre, gamma, s) < 2 * 2 * s / (N // n_lut):
                continue                               # grid rounding at discontinuities
            assert read == step_lut(pre, gamma, s), (k, j, pre, read)


def test_plaintext_relu_ramp_and_noise_model():
    from nn_tfhe.plaintext import lut_index_noise, relu_ramp
    import parameters_crypto as PC
    p = np.array([-7.0, -5.0, -4.5, -1.0, 0.0, 0.5, 3.99, 4.5, 7.9])
    got = relu_ramp(p, 4.0, 1024)
    np.testing.assert_allclose(got, [-1.0, -3.0, -3.5, 0.0, 0.0, 0.5, 3.98828125, 0.0, 0.0])
    # reference parameters: modulus switch (235 uniform terms) and key switch
    # (sqrt(2048*20) * 1 * 2^48.6 / 2^54 entries at n_lut 4)
    sig = lut_index_noise(PC.dict_params_lut, PC.dict_params_ks_LWE, PC.degree, PC.q, 4)
    ms = np.sqrt(235 / 12)
    ks = np.sqrt(2048 * 20) * 2.0 ** 48.6 / 2.0 ** 54
    assert abs(sig - np.sqrt(ms ** 2 + ks ** 2)) < 1e-9
