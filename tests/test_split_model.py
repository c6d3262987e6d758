"""CPU pins of the split-form method (tests/split_model.py restates, in the
kernels' data layout, what csrc/split_br.hip computes): the Nussbaumer
transform pair multiplies exactly, the byte limbs recombine, the Toeplitz
rows hold the negacyclic product, and one full blind rotation equals the
int128 oracle.  Planted mistakes in the model (a wrong twiddle sign, a
dropped limb pair) must fail the same checks."""
import numpy as np
import pytest

from oracle import exact as X
from oracle import tfhe_oracle as O
from tests import split_model as S


@pytest.mark.parametrize("N", [512, 1024, 2048])
def test_transform_pair_gives_2r_times_the_product(N):
    M, R = S.ring_shape(N)
    rng = np.random.default_rng(N)
    a = rng.integers(-2 ** 31, 2 ** 31, N)
    d = rng.integers(-64, 64, N)
    fa, fd = S.forward(a, M, R), S.forward(d, M, R)
    prod = np.stack([X.negacyclic_mul(fa[j], fd[j], 64) for j in range(2 * R)])
    got = S.inverse_fold(prod, M, R)
    np.testing.assert_array_equal(got, X.negacyclic_mul(a, d, 64) * (2 * R))
    # digit bins stay within two signed bytes
    assert np.abs(fd).max() <= R * 64


def test_byte_limbs_recombine_mod_2_40():
    rng = np.random.default_rng(1)
    v = rng.integers(-2 ** 62, 2 ** 62, 5000)
    limbs = S.byte_limbs(v)
    assert limbs.min() >= -128 and limbs.max() <= 127
    back = sum(limbs[:, t].astype(object) << (8 * t) for t in range(S.LIMBS))
    assert all((int(b) - int(x)) % (1 << 40) == 0 for b, x in zip(back, v))


def test_toeplitz_rows_hold_the_negacyclic_product():
    M = 64
    rng = np.random.default_rng(2)
    k = rng.integers(-2 ** 36, 2 ** 36, M)
    x = rng.integers(-2048, 2048, M)
    rev = S.rev_vectors(k[None], M)[0]                          # (LIMBS, 2M)
    T = S.toeplitz_rows(rev, M)                                 # (LIMBS, M, M)
    got = sum((T[t] @ x).astype(object) << (8 * t) for t in range(S.LIMBS))
    want = X.negacyclic_mul(k, x, 64)
    assert all((int(g) - int(w)) % (1 << 40) == 0 for g, w in zip(got, want))


def _case(N, L, lb, lq, n=6, B=2, seed=0):
    rng = np.random.default_rng(seed)
    s, Sk = rng.integers(0, 2, n), rng.integers(0, 2, N)
    bsk = O.bsk_collapse_coef(rng, s, Sk, lq, lb, L, 3, 2.0 ** 5)
    kk = O.bsk_coef_kernel_layout(bsk, n // 3, 3)
    ct = O.uniform_torus(rng, (B, 2, N), lq)
    idx = O.rotation_indices(O.uniform_torus(rng, (B, n), lq), lq, N, 3, 1)
    return ct, kk, idx


@pytest.mark.parametrize("N,L,lb,lq", [(512, 1, 7, 32), (512, 4, 7, 32), (1024, 2, 6, 24),
                                       (512, 3, 6, 20)])
def test_model_blind_rotation_equals_exact(N, L, lb, lq):
    ct, kk, idx = _case(N, L, lb, lq, seed=N + L)
    np.testing.assert_array_equal(S.blind_rotate(ct, kk, idx, lb, lq, L),
                                  X.blind_rotate(ct, kk, idx, lb, lq, L))


def test_planted_mistakes_fail(monkeypatch):
    ct, kk, idx = _case(512, 2, 7, 32, n=3, B=1, seed=4)
    want = X.blind_rotate(ct, kk, idx, 7, 32, 2)
    # a twiddle with the wrong sign in the forward transform
    real = S.mul_y
    monkeypatch.setattr(S, "mul_y", lambda x, s, M: real(x, -s if s % M else s, M))
    assert not np.array_equal(S.blind_rotate(ct, kk, idx, 7, 32, 2), want)
    monkeypatch.setattr(S, "mul_y", real)
    # a dropped limb (the top key limb cut to zero)
    real_limbs = S.byte_limbs
    monkeypatch.setattr(S, "byte_limbs", lambda v, nbits=40: real_limbs(v, 32))
    assert not np.array_equal(S.blind_rotate(ct, kk, idx, 7, 32, 2), want)
