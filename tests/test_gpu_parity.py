"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: bit-exact for integer work (rotation indices, rotate/extract, key
switch); for the FP64 blind rotation, every output coefficient within a
stated distance of the exact int128 oracle and decrypted messages identical
to the exact oracle's.
"""
import numpy as np
import pytest
import torch

from oracle import exact as X
from oracle import tfhe_oracle as O
from tests.tfhe_cases import decrypt_outputs, expected_mod, lut_at_switched_phase, make_case

pytestmark = pytest.mark.gpu

# FP64 blind-rotation tolerance: max |GPU - exact| per coefficient after the
# whole loop, in integer units mod q.  The FP64 transform's own rounding gives
# O(1) differences; the phase noise is ~2^20 and more.
BR_COEF_TOL = 64


def dev(x, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda:0")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def mult():
    from cuTFHE import multiply
    return multiply


def test_library_is_native(cuda_device):
    from cuTFHE import _core
    assert _core.lib().bslstm_abi_version() == 2


def test_rotation_indices_bit_exact(cuda_device, mult):
    rng = np.random.default_rng(11)
    for (N, n, c, n_lut) in [(2048, 750, 3, 4), (1024, 630, 3, 1), (512, 24, 2, 2), (256, 12, 4, 3)]:
        q = 1 << 32
        pk = rng.integers(-q // 2, q // 2, size=(64, n)).astype(np.float64)
        # built ties: odd multiples of q*n_lut/(4N) hit x.5 exactly
        unit = q * n_lut / (4 * N)
        if unit == int(unit):
            pk[:8] = (2 * rng.integers(-2000, 2000, size=(8, n)) + 1) * unit
        ref = O.rotation_indices(pk, 32, N, c, n_lut)
        got = host(mult.rotation_indices(dev(pk), 32, N, c, n_lut))
        assert got.dtype == np.int32
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("form", ["plain", "batched", "rows"])
@pytest.mark.parametrize("N,L,log_beta", [(128, 1, 10), (128, 4, 7), (256, 2, 8), (512, 3, 7),
                                          (1024, 3, 7), (2048, 3, 7), (2048, 4, 6), (4096, 2, 9)])
def test_mul_bootstrap_vs_exact_int128(cuda_device, mult, N, L, log_beta, form):
    case = make_case(N=N, n=12, L=L, log_beta=log_beta, collapse=3, B=7, seed=N + L)
    idx = O.rotation_indices(case.a, 32, N, 3, 1)
    ct = np.broadcast_to(case.lut, (7, 2, N))
    got = host(mult.mul_bootstrap(dev(ct), dev(case.bsk_fft_k), None, dev(idx), log_beta, 32, L,
                                  form=form))
    exact = X.blind_rotate(case.lut, case.bsk_coef_k, idx, log_beta, 32, L)
    d = O.centered_mod(got - exact, 32)
    print(f"N={N} L={L} {form}: max|gpu-exact|={np.abs(d).max()}")
    assert np.abs(d).max() <= BR_COEF_TOL


@pytest.mark.parametrize("log_q,L,lb,tol", [(36, 2, 8, 64 * 16), (24, 2, 6, 64), (20, 1, 10, 64)])
@pytest.mark.parametrize("form", ["plain", "batched", "rows"])
def test_mul_bootstrap_other_moduli(cuda_device, mult, log_q, L, lb, tol, form):
    """log2 q > 32 runs the 64-bit accumulator, log2 q < 32 the 32-bit one;
    the FP64 deviation from the exact oracle grows with q."""
    case = make_case(N=256, n=12, L=L, log_beta=lb, collapse=3, B=5, seed=log_q, log_q=log_q,
                     sigma_bsk=2.0 ** (log_q - 26), sigma_lwe=2.0 ** (log_q - 15))
    idx = O.rotation_indices(case.a, log_q, 256, 3, 1)
    ct = np.broadcast_to(case.lut, (5, 2, 256))
    got = host(mult.mul_bootstrap(dev(ct), dev(case.bsk_fft_k), None, dev(idx), lb, log_q, L,
                                  form=form))
    exact = X.blind_rotate(case.lut, case.bsk_coef_k, idx, lb, log_q, L)
    d = O.centered_mod(got - exact, log_q)
    print(f"log_q={log_q}: max|gpu-exact|={np.abs(d).max()}")
    assert np.abs(d).max() <= tol
    half = 1 << (log_q - 1)
    assert got.min() >= -half and got.max() < half


@pytest.mark.parametrize("N,L,c", [(2048, 3, 3), (1024, 2, 2), (512, 4, 1), (256, 3, 4),
                                   (512, 1, 3)])
def test_all_forms_bit_identical(cuda_device, mult, N, L, c):
    """plain, batched and rows (and pair, at N = 2048 and l <= 3) give
    identical outputs.  Odd batch (a tail workgroup with one ciphertext) and
    per-ciphertext test polynomials; collapse 1..4 (m = 2..16 patterns)."""
    case = make_case(N=N, n=12 * c, L=L, log_beta=7, collapse=c, B=9, seed=3 * N + c)
    rng = np.random.default_rng(c)
    cts = rng.integers(-2 ** 31, 2 ** 31, size=(9, 2, N))
    idx = dev(O.rotation_indices(case.a, 32, N, c, 1))
    bsk = dev(case.bsk_fft_k)
    r_plain = host(mult.mul_bootstrap(dev(cts), bsk, None, idx, 7, 32, L, form="plain"))
    r_batch = host(mult.mul_bootstrap(dev(cts), bsk, None, idx, 7, 32, L, form="batched"))
    r_rows = host(mult.mul_bootstrap(dev(cts), bsk, None, idx, 7, 32, L, form="rows"))
    np.testing.assert_array_equal(r_plain, r_batch)
    np.testing.assert_array_equal(r_plain, r_rows)
    if mult.pair_applicable(N, L, 9, 0):
        r_pair = host(mult.mul_bootstrap(dev(cts), bsk, None, idx, 7, 32, L, form="pair"))
        assert mult.pair_status() & 1 == 0
        np.testing.assert_array_equal(r_plain, r_pair)
    exact = X.blind_rotate(cts, case.bsk_coef_k, host(idx), 7, 32, L)
    assert np.abs(O.centered_mod(r_batch - exact, 32)).max() <= BR_COEF_TOL


def test_mul_bootstrap_key_slots_all_forms(cuda_device, mult):
    """Two clients' keys in one pool: every ciphertext through its slot
    equals the single-key launch with that client's key, in every form."""
    c1 = make_case(N=512, n=24, L=3, log_beta=7, B=6, seed=31)
    c2 = make_case(N=512, n=24, L=3, log_beta=7, B=6, seed=32)
    rng = np.random.default_rng(33)
    cts = dev(rng.integers(-2 ** 31, 2 ** 31, size=(6, 2, 512)))
    idx = dev(rng.integers(0, 1024, size=(6, 8, 8)).astype(np.int32))
    keys = [dev(c.bsk_fft_k) for c in (c1, c2)]
    pool = torch.stack(keys)
    slots = [0, 0, 1, 1, 0, 0]                  # pairs share a slot (batched form)
    alone = [host(mult.mul_bootstrap(cts, k, None, idx, 7, 32, 3, form="plain")) for k in keys]
    want = np.stack([alone[s][i] for i, s in enumerate(slots)])
    for form in ("plain", "batched", "rows"):
        got = host(mult.mul_bootstrap(cts, pool, None, idx, 7, 32, 3, form=form,
                                      slots=dev(np.array(slots, np.int32))))
        np.testing.assert_array_equal(got, want)


def test_mul_bootstrap_broadcast_equals_batched(cuda_device, mult):
    case = make_case(N=1024, n=24, L=3, log_beta=7, B=8, seed=5)
    idx = dev(O.rotation_indices(case.a, 32, 1024, 3, 1))
    bsk = dev(case.bsk_fft_k)
    one = host(mult.mul_bootstrap(dev(case.lut), bsk, None, idx, 7, 32, 3))
    many = host(mult.mul_bootstrap(dev(np.broadcast_to(case.lut, (8, 2, 1024))), bsk, None, idx, 7, 32, 3))
    np.testing.assert_array_equal(one, many)


def test_mul_bootstrap_deterministic(cuda_device, mult):
    case = make_case(N=2048, n=24, L=3, log_beta=7, B=4, seed=9)
    idx = dev(O.rotation_indices(case.a, 32, 2048, 3, 1))
    bsk = dev(case.bsk_fft_k)
    r1 = host(mult.mul_bootstrap(dev(case.lut), bsk, None, idx, 7, 32, 3))
    r2 = host(mult.mul_bootstrap(dev(case.lut), bsk, None, idx, 7, 32, 3))
    np.testing.assert_array_equal(r1, r2)


def test_mul_bootstrap_vs_fp64_restatement(cuda_device, mult):
    """Same algorithm as the reference kernel (FP64 FFT domain), batch of 32."""
    case = make_case(N=1024, n=30, L=3, log_beta=7, B=32, seed=21)
    idx = O.rotation_indices(case.a, 32, 1024, 3, 1)
    ct = np.broadcast_to(case.lut, (32, 2, 1024))
    got = host(mult.mul_bootstrap(dev(ct), dev(case.bsk_fft_k), None, dev(idx), 7, 32, 3))
    ref = O.blind_rotate_f64(ct, case.bsk_fft_k, idx, 7, 32, 3)
    assert np.abs(O.centered_mod(got - ref, 32)).max() <= BR_COEF_TOL


def test_rotate_extract_bit_exact(cuda_device):
    from schemas.RLWE_jax import rotate_extract
    rng = np.random.default_rng(3)
    for N, n_lut in [(1024, 1), (2048, 4), (256, 3)]:
        br = rng.integers(-2 ** 31, 2 ** 31, size=(33, 2, N))
        b = rng.integers(-2 ** 31, 2 ** 31, size=33).astype(np.float64)
        b[:4] = [0, -2 ** 31, 2 ** 31 - 1, 2 ** 20 * n_lut + 2 ** 19 * n_lut]   # incl. a tie
        masks, bodies = rotate_extract(dev(br), dev(b), 32, n_lut)
        rm, rb = O.cuboot_merge_tail(br, b, 32, n_lut)
        np.testing.assert_array_equal(host(masks).astype(np.int64), rm)
        np.testing.assert_array_equal(host(bodies).astype(np.int64), rb)


def test_key_switch_bit_exact(cuda_device):
    """The FP64 key switch (K split over workgroups, exact FP64 atomics)
    equals the oracle's float-rule key switch; 70 ciphertexts and three
    shapes, so the split, the row tiles and a short last chunk all occur."""
    from schemas.LWE_jax import key_switch_LWE
    rng = np.random.default_rng(4)
    for (n_in, n_out, lb, l) in [(1024, 630, 4, 5), (2048, 750, 4, 5), (300, 37, 2, 12)]:
        s_old = rng.integers(0, 2, n_in)
        s_new = rng.integers(0, 2, n_out)
        ka, kb = O.keyswitch_key(rng, s_new, s_old, 32, lb, l, 2.0 ** 12)
        a = rng.integers(-2 ** 31, 2 ** 31, size=(70, n_in))
        a[0, :3] = [2 ** 31 - 1, -2 ** 31, 2 ** (32 - lb - 1)]     # extremes and a tie
        # representations outside [-q/2, q/2): the float rule decomposes
        # a/q as given (a first digit of +beta/2 at +q/2, larger beyond)
        a[1, :4] = [2 ** 31, 3 * 2 ** 30, -3 * 2 ** 30, 2 ** 32 - 1]
        b = rng.integers(-2 ** 31, 2 ** 31, size=70)
        ra, rb = O.key_switch_lwe(ka, kb, a, b, 32, lb, l)
        params = {"beta_ks": 2 ** lb, "l_ks": l, "q": 2.0 ** 32}
        ga, gb = key_switch_LWE((dev(ka, torch.float64), dev(kb[..., None], torch.float64)),
                                (dev(a, torch.float64), dev(b[:, None], torch.float64)), params)
        assert O.eq_mod(host(ga).astype(np.int64), ra, 32)
        assert O.eq_mod(host(gb).astype(np.int64).reshape(-1), rb, 32)


def test_errors_are_invalid_argument(cuda_device, mult):
    from cuTFHE._core import TFHEInvalidArgument
    case = make_case(N=128, n=12, L=2, log_beta=7, B=2, seed=1)
    idx = dev(O.rotation_indices(case.a, 32, 128, 3, 1))
    with pytest.raises(TFHEInvalidArgument, match="Unsupported l"):
        mult.mul_bootstrap(dev(case.lut), dev(case.bsk_fft_k), None, idx, 7, 32, 5)
    with pytest.raises(TFHEInvalidArgument):
        mult.mul_bootstrap(dev(np.zeros((2, 100), np.int64)), dev(case.bsk_fft_k), None, idx, 7, 32, 2)
    with pytest.raises(TFHEInvalidArgument):   # wrong index count
        mult.mul_bootstrap(dev(case.lut), dev(case.bsk_fft_k), None, idx[:1, :2], 7, 32, 2)


def _pbs_gpu(case, lut_batched=True, B=None):
    from cuTFHE.multiply import log2_int  # noqa: F401
    from schemas.RLWE_jax import cuboot_merge
    Bn = len(case.b) if B is None else B
    lut = np.broadcast_to(case.lut, (Bn, 2, case.N))
    masks, bodies = cuboot_merge((dev(case.a[:Bn], torch.float64), dev(case.b[:Bn, None], torch.float64)),
                                 [dev(lut[:, 0], torch.float64), dev(lut[:, 1], torch.float64)],
                                 dev(case.bsk_ordered), 2.0 ** case.log_q, 2.0 ** case.log_beta, case.L,
                                 case.N, case.collapse, None, case.n_lut)
    return host(masks), host(bodies)


def _pbs_exact(case, B):
    idx = O.rotation_indices(case.a[:B], case.log_q, case.N, case.collapse, case.n_lut)
    br = X.blind_rotate(case.lut, case.bsk_coef_k, idx, case.log_beta, case.log_q, case.L)
    return O.cuboot_merge_tail(br, case.b[:B], case.log_q, case.n_lut)


def test_pbs_identity_lut_exhaustive_small(cuda_device):
    """Every 3-bit message through the identity LUT (noise-free LWE)."""
    case = make_case(N=512, n=36, L=3, log_beta=7, B=64, msg_bits=3, seed=2, sigma_lwe=0.0,
                     sigma_bsk=2.0 ** 6)
    masks, bodies = _pbs_gpu(case)
    dec, err = decrypt_outputs(case, masks, bodies[..., 0])
    np.testing.assert_array_equal(dec, expected_mod(case))
    assert set(case.msg.tolist()) == set(range(8))


def test_pbs_n1024_n630_identity_then_key_switch_vs_exact(cuda_device):
    """N=1024, n=630, l=3, log2 Bg=7, identity LUT, 3-bit messages (the
    reference's CPU-runnable case), then key switch back to n with base 2^4,
    l=5."""
    from schemas.LWE_jax import key_switch_LWE
    case = make_case(N=1024, n=630, L=3, log_beta=7, B=8, msg_bits=3, seed=7, sigma_bsk=2.0 ** 7)
    masks, bodies = _pbs_gpu(case)
    em, eb = _pbs_exact(case, 8)
    dec_g, err_g = decrypt_outputs(case, masks, bodies[..., 0])
    dec_x, err_x = decrypt_outputs(case, em, eb)
    np.testing.assert_array_equal(dec_g, dec_x)
    np.testing.assert_array_equal(dec_g, expected_mod(case))
    dev_max = np.abs(O.centered_mod(masks.astype(np.int64) - em, 32)).max()
    print(f"N=1024 n=630: max|gpu-exact| mask coef {dev_max}; err std gpu {err_g.std():.4g} "
          f"exact {err_x.std():.4g}")
    # key switch N=1024 -> n=630 and decrypt under s
    rng = np.random.default_rng(70)
    ka, kb = O.keyswitch_key(rng, case.s, case.S, 32, 4, 5, 2.0 ** 10)
    params = {"beta_ks": 16, "l_ks": 5, "q": 2.0 ** 32}
    ga, gb = key_switch_LWE((dev(ka, torch.float64), dev(kb[..., None], torch.float64)),
                            (dev(masks[:, 0]), dev(bodies[:, 0])), params)
    ga, gb = host(ga).astype(np.int64), host(gb).astype(np.int64).reshape(-1)
    ra, rb = O.key_switch_lwe(ka, kb, masks[:, 0].astype(np.int64), bodies[:, 0, 0].astype(np.int64), 32, 4, 5)
    assert O.eq_mod(ga, ra, 32) and O.eq_mod(gb, rb, 32)
    ph = O.centered_mod(gb - ga @ case.s, 32)
    np.testing.assert_array_equal(np.mod(np.rint(ph / case.delta_out).astype(np.int64), 16),
                                  case.msg)


def test_pbs_n2048_n750_many_lut_vs_exact(cuda_device):
    """N=2048, n=750, l=3, log2 Bg=7, 4-bit signed messages, 4 LUTs per test
    polynomial: decryptions identical to the exact int128 oracle's."""
    case = make_case(N=2048, n=750, L=3, log_beta=7, B=6, n_lut=4, msg_bits=4, seed=8,
                     sigma_bsk=2.0 ** 5, kind="many")
    masks, bodies = _pbs_gpu(case)
    em, eb = _pbs_exact(case, 6)
    dec_g, err_g = decrypt_outputs(case, masks, bodies[..., 0])
    dec_x, err_x = decrypt_outputs(case, em, eb)
    np.testing.assert_array_equal(dec_g, dec_x)
    d = np.abs(O.centered_mod(masks.astype(np.int64) - em, 32)).max()
    print(f"N=2048 n=750: max|gpu-exact| {d}; err std gpu {err_g.std():.4g} exact {err_x.std():.4g}")
    assert d <= BR_COEF_TOL
    # both carry exactly the test polynomial at the modulus-switched phase
    want = lut_at_switched_phase(case, 6)
    np.testing.assert_array_equal(dec_x, want)
    np.testing.assert_array_equal(dec_g, want)


def test_multiply_seq_monomial_dropin(cuda_device, mult):
    """Float64 surface with the reference-layout BSK (2,2,l,n_eq,N/2)."""
    case = make_case(N=256, n=24, L=2, log_beta=9, B=5, seed=12)
    pk = dev(case.a, torch.float64)
    ct = dev(np.broadcast_to(case.lut, (5, 2, 256)), torch.float64)
    rot = dev(np.stack([O.negacyclic_fft(np.eye(256)[r % 256] * (-1) ** (r // 256)) for r in range(512)]))
    out = host(mult.multiply_seq_monomial(ct, pk, dev(case.bsk_ordered), 32, 9, 2, 256, 3, rot, 1))
    assert out.dtype == np.float64 and out.shape == (5, 2, 256)
    idx = O.rotation_indices(case.a, 32, 256, 3, 1)
    exact = X.blind_rotate(case.lut, case.bsk_coef_k, idx, 9, 32, 2)
    assert np.abs(O.centered_mod(out.astype(np.int64) - exact, 32)).max() <= BR_COEF_TOL


def test_heaviside_bootstrapper_cuboot(cuda_device):
    """KS (N -> n) then many-LUT PBS, as HeavysideBoostrapper.cuboot does
    (nn_tfhe/activation.py:148-162)."""
    from nn_tfhe.activation import HeavysideBoostrapper
    case = make_case(N=512, n=48, L=3, log_beta=7, B=16, n_lut=2, msg_bits=3, seed=31,
                     sigma_bsk=2.0 ** 5, kind="many")
    rng = np.random.default_rng(31)
    # inputs encrypted under the GLWE key S (dimension N), as in the LSTM
    a_big, b_big = O.lwe_encrypt(rng, case.msg * case.delta_in, case.S, 32, 2.0 ** 12)
    ka, kb = O.keyswitch_key(rng, case.s, case.S, 32, 4, 6, 2.0 ** 8)
    p_lut = {"q": 2.0 ** 32, "beta_bs": 2 ** 7, "l_bs": 3, "degree": 512}
    p_ks = {"q": 2.0 ** 32, "beta_ks": 16, "l_ks": 6, "degree": 48}
    bs = HeavysideBoostrapper(torch.as_tensor(case.bsk_ordered), (ka.astype(np.float64),
                              kb[..., None].astype(np.float64)), None, None, {}, p_lut, p_ks, {},
                              3, None, 2)
    lut = np.broadcast_to(case.lut, (16, 2, 512))
    masks, bodies = bs.cuboot((dev(a_big, torch.float64), dev(b_big[:, None], torch.float64)),
                              [dev(lut[:, 0], torch.float64), dev(lut[:, 1], torch.float64)])
    dec, _ = decrypt_outputs(case, host(masks), host(bodies)[..., 0])
    np.testing.assert_array_equal(dec, expected_mod(case))


def test_br_and_bootstrapping_vs_exact(cuda_device):
    """BR() (RLWE_jax.py:305-315: blind rotation, then X^(-b_ms)) against the
    int128 oracle, and bootstrapping() (RLWE_jax.py:234-274), batched and on
    one ciphertext, decrypting to the identity LUT."""
    from schemas.RLWE_jax import BR, bootstrapping
    case = make_case(N=512, n=48, L=3, log_beta=7, B=12, msg_bits=3, seed=41, sigma_bsk=2.0 ** 5)
    B, N, q = 12, case.N, 2.0 ** 32
    lut = np.broadcast_to(case.lut, (B, 2, N))
    pk, b = dev(case.a, torch.float64), dev(case.b[:, None], torch.float64)
    lut_rows = [dev(lut[:, 0], torch.float64), dev(lut[:, 1], torch.float64)]
    bsk = dev(case.bsk_ordered)
    br = host(BR((pk, b), lut_rows, bsk, q, 2.0 ** 7, 3, N, 3, None, 1))
    assert br.shape == (B, 2, N) and br.dtype == np.float64
    idx = O.rotation_indices(case.a, 32, N, 3, 1)
    exact = X.blind_rotate(case.lut, case.bsk_coef_k, idx, 7, 32, 3)
    b_ms = np.rint(2 * N * case.b.astype(np.float64) / q).astype(np.int64)
    for i in range(B):
        for c in range(2):
            want = O.coef_rotation(exact[i, c], -b_ms[i])
            assert np.abs(O.centered_mod(br[i, c].astype(np.int64) - want, 32)).max() <= BR_COEF_TOL
    masks, bodies = bootstrapping((pk, b), lut_rows, bsk, q, 2.0 ** 7, 3, N, 3)
    masks, bodies = host(masks), host(bodies)
    assert masks.shape == (B, N) and bodies.shape == (B, 1)
    dec, _ = decrypt_outputs(case, masks[:, None], bodies)
    np.testing.assert_array_equal(dec, expected_mod(case))
    # bootstrapping()'s body is coefficient 0 of BR()'s body row (mod q: the
    # extraction reduces, BR's float rotation does not)
    assert (O.centered_mod(bodies[:, 0].astype(np.int64) - br[:, 1, 0].astype(np.int64), 32) == 0).all()
    m1, b1 = bootstrapping((pk[0], b[0]), [lut_rows[0][0], lut_rows[1][0]], bsk, q, 2.0 ** 7, 3, N, 3)
    np.testing.assert_array_equal(host(m1), masks[0])
    np.testing.assert_array_equal(host(b1).reshape(-1), bodies[0])


def test_heaviside_bootstrapper_call(cuda_device):
    """HeavysideBoostrapper.__call__ (activation.py:120-146): key switch N -> n,
    then bootstrapping() with one LUT; decrypts to LUT(m) and equals cuboot
    with n_lut = 1."""
    from nn_tfhe.activation import HeavysideBoostrapper
    case = make_case(N=512, n=48, L=3, log_beta=7, B=16, msg_bits=3, seed=43, sigma_bsk=2.0 ** 5)
    rng = np.random.default_rng(43)
    a_big, b_big = O.lwe_encrypt(rng, case.msg * case.delta_in, case.S, 32, 2.0 ** 12)
    ka, kb = O.keyswitch_key(rng, case.s, case.S, 32, 4, 6, 2.0 ** 8)
    p_lut = {"q": 2.0 ** 32, "beta_bs": 2 ** 7, "l_bs": 3, "degree": 512}
    p_ks = {"q": 2.0 ** 32, "beta_ks": 16, "l_ks": 6, "degree": 48}
    bs = HeavysideBoostrapper(torch.as_tensor(case.bsk_ordered), (ka.astype(np.float64),
                              kb[..., None].astype(np.float64)), None, None, {}, p_lut, p_ks, {},
                              3, None, 1)
    lut = np.broadcast_to(case.lut, (16, 2, 512))
    h = (dev(a_big, torch.float64), dev(b_big[:, None], torch.float64))
    lut_rows = [dev(lut[:, 0], torch.float64), dev(lut[:, 1], torch.float64)]
    masks, bodies = bs(h, lut_rows)
    masks, bodies = host(masks), host(bodies)
    dec, _ = decrypt_outputs(case, masks[:, None], bodies)
    np.testing.assert_array_equal(dec, expected_mod(case))
    cm, cb = bs.cuboot(h, lut_rows)
    np.testing.assert_array_equal(host(cm)[:, 0], masks)
    np.testing.assert_array_equal(host(cb)[:, 0, 0], bodies[:, 0])


@pytest.mark.parametrize("dtype", [torch.float64, torch.int64])
def test_text_rotation_and_extraction_on_hip_kernels(cuda_device, dtype):
    """text_jax.sample_extract / rotate_ciphertext and polynomial_jax.coef_rotation
    (text_jax.py:191-208,348-362; polynomial_jax.py:150-169) through the HIP
    kernels, float64 and int64 carriers, against the oracle."""
    from schemas.polynomial_jax import coef_rotation
    from schemas.text_jax import rotate_ciphertext, sample_extract
    rng = np.random.default_rng(44)
    N, R = 256, 6
    A = rng.integers(-2 ** 31, 2 ** 31, (R, N))
    Bq = rng.integers(-2 ** 31, 2 ** 31, (R, N))
    e = rng.integers(-3 * N, 3 * N, R)
    ta, tb = dev(A, dtype), dev(Bq, dtype)
    got = host(coef_rotation(ta, dev(e, torch.int64)))
    assert got.dtype == (np.float64 if dtype == torch.float64 else np.int64)
    for i in range(R):
        np.testing.assert_array_equal(got[i].astype(np.int64), O.coef_rotation(A[i], e[i]))
    ra, rb = rotate_ciphertext((ta, tb), dev(e, torch.int64))
    ra, rb = host(ra), host(rb)
    for i in range(R):
        np.testing.assert_array_equal(ra[i].astype(np.int64), O.coef_rotation(A[i], e[i]))
        np.testing.assert_array_equal(rb[i].astype(np.int64), O.coef_rotation(Bq[i], e[i]))
    for alpha in (0, 1, 77, N - 1):
        ma, mb = sample_extract((ta, tb), alpha)
        ma, mb = host(ma), host(mb)
        assert ma.shape == (R, N) and mb.shape == (R, 1)
        for i in range(R):
            wa, wb = O.sample_extract((A[i], Bq[i]), alpha)
            np.testing.assert_array_equal(ma[i].astype(np.int64), np.asarray(wa, np.int64))
            assert int(mb[i, 0]) == int(np.asarray(wb).reshape(-1)[0])


def test_pbs_config2_full_batch(cuda_device, mult):
    """The batched-PBS workload at its full size: 16384 ciphertexts, n=750,
    N=2048, l=3, log2 Bg=7, 4 LUTs, 4-bit messages.  Size-independent
    properties: the plain, batched and rows launch forms agree bit for bit on
    every ciphertext (16384 ciphertexts exceed the pair form's one pair per CU),
    every one of the 65536 outputs decrypts to the test polynomial at its
    modulus-switched phase (an exact expectation, no tolerance), and a sample
    equals the int128 oracle within the FP64 bound."""
    from schemas.RLWE_jax import rotate_extract
    B = 16384
    case = make_case(N=2048, n=750, L=3, log_beta=7, B=B, n_lut=4, msg_bits=4, seed=16,
                     sigma_bsk=2.0 ** 5, kind="many")
    idx = mult.rotation_indices(dev(case.a, torch.float64), 32, 2048, 3, 4)
    np.testing.assert_array_equal(host(idx), O.rotation_indices(case.a, 32, 2048, 3, 4))
    lut = dev(case.lut)                          # one test polynomial for the whole batch
    bsk = dev(case.bsk_fft_k)
    outs = {f: mult.mul_bootstrap(lut, bsk, None, idx, 7, 32, 3, form=f)
            for f in ("batched", "plain", "rows")}
    for f in ("plain", "rows"):
        assert torch.equal(outs[f], outs["batched"]), f
        del outs[f]
    masks, bodies = rotate_extract(outs["batched"], dev(case.b, torch.float64), 32, 4)
    del outs
    dec, _ = decrypt_outputs(case, host(masks), host(bodies))
    np.testing.assert_array_equal(dec, lut_at_switched_phase(case))
    print(f"B={B}: decryptions equal LUT(m) on {(dec == expected_mod(case)).mean():.5f} "
          f"of the outputs (the rest are modulus-switch misses)")
    sample = [0, 5461, 16383]
    br = host(mult.mul_bootstrap(lut, bsk, None, idx[sample].contiguous(), 7, 32, 3, form="batched"))
    exact = X.blind_rotate(case.lut, case.bsk_coef_k, O.rotation_indices(case.a[sample], 32, 2048, 3, 4),
                           7, 32, 3)
    assert np.abs(O.centered_mod(br - exact, 32)).max() <= BR_COEF_TOL
