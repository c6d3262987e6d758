"""Parity of the split-form blind rotation (exact int8-limb MFMA GEMMs,
bslstm_mul_bootstrap_split through the C ABI) with the exact int128 oracle.

Bar: bit-exact.  The split form computes every negacyclic product exactly,
so its output must equal the int128 oracle's coefficient for coefficient at
every degree, gadget depth and modulus it accepts.  Ciphertexts are
independent, so the oracle (O(N^2) per product) checks a sample of each batch
-- the first, one in the middle, the last -- and the rest of the batch must
equal the FP64 batched form within the FP64 bar (which the oracle pins
separately in test_gpu_parity.py).
"""
import numpy as np
import pytest
import torch

from oracle import exact as X
from oracle import tfhe_oracle as O
from tests.tfhe_cases import decrypt_outputs, expected_mod, make_case

pytestmark = pytest.mark.gpu

FP64_TOL = 64     # the FP64 forms' bar against the exact product (test_gpu_parity.py)


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


def _split(mult, ct, bsk_fft_k, idx, log_beta, log_q, L, chunk=None):
    bsk = dev(bsk_fft_k)
    key = mult.prepare_split_key(bsk, L, ct.shape[-1])
    kw = {} if chunk is None else {"chunk": chunk}
    return host(mult.mul_bootstrap_split(dev(ct), key, bsk.shape, dev(idx, torch.int32), log_beta,
                                         log_q, L, **kw))


def _sample(B):
    return sorted({0, B // 2, B - 1})


@pytest.mark.parametrize("N,L,log_beta,log_q,n,B", [
    (512, 1, 7, 32, 12, 70), (512, 2, 7, 32, 12, 130), (512, 3, 7, 32, 12, 40),
    (512, 4, 7, 32, 12, 9), (1024, 1, 7, 32, 9, 129), (1024, 3, 7, 32, 9, 64),
    (1024, 4, 6, 24, 9, 33), (2048, 1, 7, 32, 6, 65), (2048, 2, 7, 32, 6, 17),
    (2048, 3, 7, 32, 6, 70), (2048, 4, 7, 32, 6, 5), (2048, 3, 6, 20, 6, 8),
])
def test_split_blind_rotation_equals_exact(cuda_device, mult, N, L, log_beta, log_q, n, B):
    case = make_case(N=N, n=n, L=L, log_beta=log_beta, collapse=3, B=B, seed=7 * N + L,
                     log_q=log_q)
    rng = np.random.default_rng(N + L)
    ct = O.uniform_torus(rng, (B, 2, N), log_q)          # one accumulator per ciphertext
    idx = O.rotation_indices(case.a, log_q, N, 3, 1)
    got = _split(mult, ct, case.bsk_fft_k, idx, log_beta, log_q, L)
    assert got.shape == (B, 2, N)
    sel = _sample(B)
    want = X.blind_rotate(ct[sel], case.bsk_coef_k, idx[sel], log_beta, log_q, L)
    np.testing.assert_array_equal(got[sel], want)
    # the whole batch against the FP64 batched form
    fp = host(mult.mul_bootstrap(dev(ct), dev(case.bsk_fft_k), None, dev(idx, torch.int32), log_beta,
                                 log_q, L, form="batched"))
    dev_ = np.abs(O.centered_mod(got - fp, log_q)).max()
    assert dev_ <= FP64_TOL, dev_


def test_split_chunks_and_broadcast(cuda_device, mult):
    """A batch over several chunks (chunk 128 at N = 2048) and a broadcast test
    polynomial give the same outputs as one chunk and the exact oracle."""
    N, L, lb = 2048, 3, 7
    case = make_case(N=N, n=6, L=L, log_beta=lb, collapse=3, B=300, seed=5)
    idx = O.rotation_indices(case.a, 32, N, 3, 1)
    one = _split(mult, case.lut, case.bsk_fft_k, idx, lb, 32, L)              # broadcast LUT
    many = _split(mult, case.lut, case.bsk_fft_k, idx, lb, 32, L, chunk=128)
    np.testing.assert_array_equal(one, many)
    sel = [0, 127, 128, 299]
    want = X.blind_rotate(case.lut, case.bsk_coef_k, idx[sel], lb, 32, L)
    np.testing.assert_array_equal(one[sel], want)


def test_split_pbs_decrypts_to_lut(cuda_device, mult):
    """Many-LUT PBS at the batched-throughput parameters (n = 750 scaled down
    to 30 for the oracle, N = 2048, l = 3, log2 Bg = 7, 4 LUTs): the split form
    through cuboot_merge-style rotation and extraction decrypts every output to
    LUT(m), exactly like the int128 oracle."""
    N, L, lb, n_lut = 2048, 3, 7, 4
    case = make_case(N=N, n=30, L=L, log_beta=lb, collapse=3, B=48, n_lut=n_lut, msg_bits=4,
                     seed=9, sigma_bsk=2.0 ** 5, kind="many")
    idx = O.rotation_indices(case.a, 32, N, 3, n_lut)
    br = _split(mult, case.lut, case.bsk_fft_k, idx, lb, 32, L)
    want = X.blind_rotate(case.lut, case.bsk_coef_k, idx[:4], lb, 32, L)
    np.testing.assert_array_equal(br[:4], want)
    masks, bodies = O.cuboot_merge_tail(br, case.b, 32, n_lut)
    got, _ = decrypt_outputs(case, masks, bodies)
    np.testing.assert_array_equal(got, expected_mod(case))


def test_split_form_through_mul_bootstrap(cuda_device, mult):
    """form="split" of mul_bootstrap prepares and caches the key and equals
    the exact oracle."""
    N, L, lb = 1024, 2, 7
    case = make_case(N=N, n=9, L=L, log_beta=lb, collapse=3, B=5, seed=3)
    idx = O.rotation_indices(case.a, 32, N, 3, 1)
    bsk = dev(case.bsk_fft_k)
    got = host(mult.mul_bootstrap(dev(case.lut), bsk, None, dev(idx, torch.int32), lb, 32, L,
                                  form="split"))
    again = host(mult.mul_bootstrap(dev(case.lut), bsk, None, dev(idx, torch.int32), lb, 32, L,
                                    form="split"))
    np.testing.assert_array_equal(got, again)
    np.testing.assert_array_equal(got, X.blind_rotate(case.lut, case.bsk_coef_k, idx, lb, 32, L))


def test_split_key_preparation_follows_the_key_tensor(cuda_device, mult):
    """Two different keys of one shape, each in a fresh tensor (the second may
    reuse the first's freed memory): each call prepares its own key."""
    N, L, lb = 512, 2, 7
    for seed in (21, 22):
        case = make_case(N=N, n=9, L=L, log_beta=lb, collapse=3, B=3, seed=seed)
        idx = O.rotation_indices(case.a, 32, N, 3, 1)
        bsk = dev(case.bsk_fft_k)
        got = host(mult.mul_bootstrap(dev(case.lut), bsk, None, dev(idx, torch.int32), lb, 32, L,
                                      form="split"))
        np.testing.assert_array_equal(got, X.blind_rotate(case.lut, case.bsk_coef_k, idx, lb, 32, L))
        del bsk
