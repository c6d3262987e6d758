"""The pair form of the blind rotation (two workgroups per ciphertext that
exchange their rows' digit spectra through device memory once per CMux step):
bit-identical to the one-workgroup forms, which are pinned to the int128
oracle elsewhere; the status word shows no wait that ran out; a forced tiny
wait bound still ends every kernel (the outputs are then not used) and leaves
the next call correct."""
import numpy as np
import pytest
import torch

from oracle import exact as X
from oracle import lstm_oracle as L
from oracle import tfhe_oracle as O
from tests.tfhe_cases import make_case

pytestmark = pytest.mark.gpu

BR_COEF_TOL = 64      # FP64 blind rotation vs the exact product, integer units mod q


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


def _max_batch(N=2048):
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return (cus if N == 1024 else cus // 2) // 8 * 8


@pytest.mark.parametrize("N,L_,c,B", [(2048, 3, 3, 1), (2048, 3, 3, 9), (2048, 2, 2, 40),
                                      (2048, 1, 3, 17), (2048, 3, 3, -1), (1024, 3, 3, 1),
                                      (1024, 3, 3, 57), (1024, 2, 3, 130), (1024, 3, 3, -1)])
def test_pair_equals_one_workgroup_forms_torus32(cuda_device, mult, N, L_, c, B):
    """log2 q = 32 (32-bit accumulator), per-ciphertext test polynomials,
    ragged batches (the last group of 8 pairs part-filled) and the largest
    batch the device holds (B = -1: two workgroups per CU at N = 1024, one
    at N = 2048)."""
    B = _max_batch(N) if B < 0 else B
    case = make_case(N=N, n=12 * c, L=L_, log_beta=7, collapse=c, B=B, seed=7 * B + L_)
    rng = np.random.default_rng(B + c)
    cts = dev(rng.integers(-2 ** 31, 2 ** 31, size=(B, 2, N)))
    idx = dev(O.rotation_indices(case.a, 32, N, c, 1))
    bsk = dev(case.bsk_fft_k)
    r_rows = host(mult.mul_bootstrap(cts, bsk, None, idx, 7, 32, L_, form="rows"))
    r_pair = host(mult.mul_bootstrap(cts, bsk, None, idx, 7, 32, L_, form="pair"))
    status = mult.pair_status()
    print(f"N={N} B={B} L={L_} c={c}: status word {status:#x}")
    assert status & 1 == 0
    np.testing.assert_array_equal(r_pair, r_rows)
    if B <= 9:
        exact = X.blind_rotate(host(cts), case.bsk_coef_k, host(idx), 7, 32, L_)
        assert np.abs(O.centered_mod(r_pair - exact, 32)).max() <= BR_COEF_TOL


def test_pair_broadcast_test_polynomial(cuda_device, mult):
    """One test polynomial for the whole batch (ct stride 0)."""
    case = make_case(N=2048, n=36, L=3, log_beta=7, collapse=3, B=24, seed=91)
    idx = dev(O.rotation_indices(case.a, 32, 2048, 3, 1))
    bsk = dev(case.bsk_fft_k)
    lut = dev(case.lut)
    r_plain = host(mult.mul_bootstrap(lut, bsk, None, idx, 7, 32, 3, form="plain"))
    r_pair = host(mult.mul_bootstrap(lut, bsk, None, idx, 7, 32, 3, form="pair"))
    assert mult.pair_status() & 1 == 0
    np.testing.assert_array_equal(r_pair, r_plain)


@pytest.mark.parametrize("log_q,lb,l", [(64, 13, 2), (48, 12, 3), (36, 8, 2)])
def test_pair_wide_moduli_equal_rows(cuda_device, mult, log_q, lb, l):
    """log2 q > 32: the LSTM's modulus 2^64 with l*log2_beta + 1 <= 32 (the
    top-bits accumulator), 2^48 at l*log2_beta + 1 > 32 (the 64-bit one) and
    2^36; the pair form equals the rows form, and at 2^64 the phase is near
    the exact oracle's.  That bound is 2^-16 q at N = 2048: an FP64 product
    error (~2^31 here) now and then moves a coefficient across a last-digit
    boundary (step 2^38 at beta^l = 2^26), and each such flip adds a key row
    whose phase carries 2^38 * s; over 2048 coefficients x 8 steps the flips
    add up to ~2^44-2^45 (measured 2^44.6), far below the LSTM's message
    scale 2^56."""
    rng = np.random.default_rng(log_q)
    N, n, c, B = 2048, 24, 3, 12
    s = rng.integers(0, 2, n)
    S = rng.integers(0, 2, N)
    bsk_coef = L.bsk_kernel_coef(rng, s, S, log_q, lb, l, c, 2.0 ** (log_q - 47))
    bsk = dev(O.negacyclic_fft(bsk_coef.astype(np.float64)))
    ct = dev(L.uniform(rng, (B, 2, N), log_q))
    idx = dev(rng.integers(0, 2 * N, (B, n // c, 2 ** c)).astype(np.int32))
    r_rows = host(mult.mul_bootstrap(ct, bsk, None, idx, lb, log_q, l, form="rows"))
    r_pair = host(mult.mul_bootstrap(ct, bsk, None, idx, lb, log_q, l, form="pair"))
    assert mult.pair_status() & 1 == 0
    np.testing.assert_array_equal(r_pair, r_rows)
    if log_q == 64:
        want = X.blind_rotate(host(ct)[:2], bsk_coef, host(idx)[:2], lb, 64, l)
        for i in range(2):
            ph_got = L.rlwe_phase(r_pair[i, 0], r_pair[i, 1], S, 64)
            ph_want = L.rlwe_phase(want[i, 0], want[i, 1], S, 64)
            with np.errstate(over="ignore"):   # int64 differences wrap: the class mod 2^64
                d = L.tmod(np.asarray(ph_got, np.int64) - np.asarray(ph_want, np.int64), 64)
            assert np.abs(d.astype(np.float64)).max() <= 2.0 ** -16 * 2.0 ** 64


def test_pair_key_slots(cuda_device, mult):
    """Ciphertexts of two clients in one launch, any slot pattern (the two
    workgroups of a ciphertext both read that ciphertext's slot)."""
    c1 = make_case(N=2048, n=24, L=3, log_beta=7, B=7, seed=41)
    c2 = make_case(N=2048, n=24, L=3, log_beta=7, B=7, seed=42)
    rng = np.random.default_rng(43)
    cts = dev(rng.integers(-2 ** 31, 2 ** 31, size=(7, 2, 2048)))
    idx = dev(rng.integers(0, 4096, size=(7, 8, 8)).astype(np.int32))
    keys = [dev(c.bsk_fft_k) for c in (c1, c2)]
    pool = torch.stack(keys)
    slots = [0, 1, 1, 0, 1, 0, 0]
    alone = [host(mult.mul_bootstrap(cts, k, None, idx, 7, 32, 3, form="plain")) for k in keys]
    want = np.stack([alone[s][i] for i, s in enumerate(slots)])
    got = host(mult.mul_bootstrap(cts, pool, None, idx, 7, 32, 3, form="pair",
                                  slots=dev(np.array(slots, np.int32))))
    assert mult.pair_status() & 1 == 0
    np.testing.assert_array_equal(got, want)


def test_pair_tiny_wait_bound_ends_and_next_call_is_correct(cuda_device, mult):
    """spin_limit = 1 (one poll per wait): waits may run out, which the status
    word reports; every kernel still ends and a following call with the
    default bound is exact again.  The outputs of the forced call are not
    checked."""
    case = make_case(N=2048, n=36, L=3, log_beta=7, collapse=3, B=32, seed=51)
    idx = dev(O.rotation_indices(case.a, 32, 2048, 3, 1))
    bsk = dev(case.bsk_fft_k)
    lut = dev(case.lut)
    from cuTFHE import _core
    mult.pair_errors()                               # start the error record afresh
    mult.mul_bootstrap(lut, bsk, None, idx, 7, 32, 3, form="pair", spin_limit=1)
    torch.cuda.synchronize()
    forced = mult.pair_status()
    print(f"status word after spin_limit=1: {forced:#x}")
    assert forced & ~3 == 0
    r_pair = host(mult.mul_bootstrap(lut, bsk, None, idx, 7, 32, 3, form="pair"))
    last = mult.pair_status()
    assert last & 1 == 0
    r_rows = host(mult.mul_bootstrap(lut, bsk, None, idx, 7, 32, 3, form="rows"))
    np.testing.assert_array_equal(r_pair, r_rows)
    # the record keeps the forced call's word after the clean call cleared its own
    assert mult.pair_errors(reset=False) == forced | last
    if forced & 1:
        with pytest.raises(_core.TFHEInternalError, match="partner"):
            mult.check_pair_errors()
    else:
        mult.check_pair_errors()
    assert mult.pair_errors() == 0                   # the check started the record again


def test_pair_rejects_outside_its_range(cuda_device, mult):
    case = make_case(N=512, n=24, L=3, log_beta=7, collapse=3, B=2, seed=61)
    idx = dev(O.rotation_indices(case.a, 32, 512, 3, 1))
    from cuTFHE import _core
    with pytest.raises(_core.TFHEInvalidArgument, match="degree 1024 or 2048"):
        mult.mul_bootstrap(dev(case.lut), dev(case.bsk_fft_k), None, idx, 7, 32, 3, form="pair")
    big = _max_batch() + 8
    case = make_case(N=2048, n=6, L=1, log_beta=7, collapse=3, B=big, seed=62)
    idx = dev(O.rotation_indices(case.a, 32, 2048, 3, 1))
    with pytest.raises(_core.TFHEInvalidArgument, match="at most"):
        mult.mul_bootstrap(dev(case.lut), dev(case.bsk_fft_k), None, idx, 7, 32, 1, form="pair")
    assert not mult.pair_applicable(2048, 1, big, 0)
    assert mult.pair_applicable(2048, 3, 8, 0)
    assert mult.pair_applicable(1024, 3, _max_batch(1024), 0)
    assert not mult.pair_applicable(1024, 3, _max_batch(1024) + 8, 0)
    assert not mult.pair_applicable(512, 3, 8, 0)


def test_pair_in_a_graph_replays_equal_eager(cuda_device, mult):
    """The pair form captured in a HIP graph (its workspace cleared by a fill
    kernel, not a memset): every replay, with eager work and synchronisations
    in between and new inputs copied into the captured buffers, equals the
    eager call on the same inputs."""
    N, L_, c, B = 2048, 2, 3, 16
    case = make_case(N=N, n=24, L=L_, log_beta=7, collapse=c, B=B, seed=71)
    rng = np.random.default_rng(72)
    bsk = dev(case.bsk_fft_k)
    cts = [dev(rng.integers(-2 ** 31, 2 ** 31, size=(B, 2, N))) for _ in range(3)]
    idxs = [dev(rng.integers(0, 2 * N, size=(B, 8, 8)).astype(np.int32)) for _ in range(3)]
    want = [host(mult.mul_bootstrap(ct, bsk, None, ix, 7, 32, L_, form="pair"))
            for ct, ix in zip(cts, idxs)]
    ct_buf, idx_buf = cts[0].clone(), idxs[0].clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):                   # warm-up on the capture stream
        mult.mul_bootstrap(ct_buf, bsk, None, idx_buf, 7, 32, L_, form="pair")
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = mult.mul_bootstrap(ct_buf, bsk, None, idx_buf, 7, 32, L_, form="pair")
    for rep in range(6):
        k = rep % 3
        ct_buf.copy_(cts[k])
        idx_buf.copy_(idxs[k])
        g.replay()
        got = host(out)
        np.testing.assert_array_equal(got, want[k], err_msg=f"replay {rep}")
        # eager work between replays, including another pair-form call
        host(mult.mul_bootstrap(cts[(k + 1) % 3], bsk, None, idxs[(k + 1) % 3], 7, 32, L_,
                                form="pair"))
