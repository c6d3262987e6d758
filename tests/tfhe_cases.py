"""Seeded PBS test cases built with the oracle (keys, LWE inputs, test
polynomials) and shared by the CPU and GPU tests."""
from __future__ import annotations

import functools
from dataclasses import dataclass, field

import numpy as np

from oracle import tfhe_oracle as O

# many-LUT sub-functions on signed 4-bit messages (sign, ReLU-clip, two steps)
MANY_LUT_FNS = (
    lambda x: 1 if x >= 0 else 0,
    lambda x: min(max(x, 0), 7),
    lambda x: 1 if x >= -4 else 0,
    lambda x: 1 if x >= 4 else 0,
)


@dataclass
class PBSCase:
    N: int
    n: int
    L: int
    log_beta: int
    collapse: int
    n_lut: int
    msg_bits: int
    log_q: int
    s: np.ndarray            # LWE secret (n,)
    S: np.ndarray            # GLWE secret (N,)
    bsk_coef: np.ndarray     # (2[L/R], 2[pk/b], n_eq, L, N) int64
    bsk_ordered: np.ndarray  # (2, 2, L, n_eq, N/2) complex128 (reference prepared layout)
    bsk_fft_k: np.ndarray    # (n_iter, 2, L, 2, m, N/2) complex128 (kernel layout)
    bsk_coef_k: np.ndarray   # (n_iter, 2, L, 2, m, N) int64
    a: np.ndarray            # (B, n) int64 LWE masks
    b: np.ndarray            # (B,) int64 LWE bodies
    msg: np.ndarray          # (B,) messages
    lut: np.ndarray          # (2, N) trivially encrypted test polynomial (A=0, B=T)
    delta_in: int
    delta_out: int
    fns: tuple = field(default_factory=tuple)

    @property
    def n_iter(self) -> int:
        return self.n // self.collapse

    def expected(self) -> np.ndarray:
        """(B, n_lut) expected output messages (LUT(m))."""
        if self.fns:
            return np.array([[f(int(x)) for f in self.fns] for x in self.msg], np.int64)
        return self.msg[:, None].copy()


@functools.lru_cache(maxsize=16)
def make_case(N=64, n=12, L=3, log_beta=7, collapse=3, B=16, n_lut=1, msg_bits=3, seed=0,
              sigma_bsk=2.0 ** 7, sigma_lwe=2.0 ** 17, log_q=32, kind="identity") -> PBSCase:
    rng = np.random.default_rng(seed)
    s = rng.integers(0, 2, n)
    S = rng.integers(0, 2, N)
    bsk = O.bsk_collapse_coef(rng, s, S, log_q, log_beta, L, collapse, sigma_bsk)
    n_iter = n // collapse
    ordered = O.bsk_to_fourier_ordered(bsk)
    bsk_fft_k = np.ascontiguousarray(O.reshape_bsk_for_kernel(ordered, n_iter, collapse))
    bsk_coef_k = np.ascontiguousarray(O.bsk_coef_kernel_layout(bsk, n_iter, collapse))
    delta = 1 << (log_q - msg_bits - 1)
    fns: tuple = ()
    if kind == "identity":
        # unsigned messages with a padding bit; T[j] = delta*m(j)
        msg = rng.integers(0, 1 << msg_bits, B)
        W = N >> msg_bits
        j = np.arange(N)
        T = delta * (((j + W // 2) // W) % (1 << msg_bits))
        delta_out = delta
    else:
        msg = rng.integers(-(1 << (msg_bits - 1)), 1 << (msg_bits - 1), B)
        fns = MANY_LUT_FNS[:n_lut]
        delta_out = delta
        T = O.lut_from_function(N, n_lut, fns, msg_bits, delta_out)
    a, b = O.lwe_encrypt(rng, msg * delta, s, log_q, sigma_lwe)
    lut = np.stack([np.zeros(N, np.int64), np.asarray(T, np.int64)])
    return PBSCase(N, n, L, log_beta, collapse, n_lut, msg_bits, log_q, s, S, bsk, ordered,
                   bsk_fft_k, bsk_coef_k, a, b, msg, lut, delta, delta_out, fns)


def decrypt_outputs(case: PBSCase, masks: np.ndarray, bodies: np.ndarray):
    """Decrypt extracted LWEs (B, n_lut, N)/(B, n_lut) under S.

    Returns (messages (B, n_lut), phase errors (B, n_lut) vs the ideal
    delta_out * expected)."""
    masks = np.asarray(np.rint(masks), np.int64)
    bodies = np.asarray(np.rint(bodies), np.int64).reshape(masks.shape[0], -1)
    ph = O.centered_mod(bodies - masks @ case.S.astype(np.int64), case.log_q)
    mod = 1 << (case.msg_bits + 1)
    dec = np.mod(np.rint(ph / case.delta_out).astype(np.int64), mod)
    exp = case.expected()[: masks.shape[0]]
    err = O.centered_mod(ph - exp * case.delta_out, case.log_q)
    return dec, err


def expected_mod(case: PBSCase, B: int | None = None) -> np.ndarray:
    return np.mod(case.expected()[:B], 1 << (case.msg_bits + 1))


def switched_rotation(case: PBSCase, B: int | None = None) -> np.ndarray:
    """(B,) exponent e of the noise-free programmable bootstrap: the blind
    rotation multiplies the test polynomial by X^idx[g, p_g] for every group g
    of c secret bits (p_g = the group's bit pattern, most significant first,
    as all_binary_vectors orders them), then the extraction rotates by
    -b_ms: e = sum_g idx[g, p_g] - b_ms mod 2N."""
    Bn = len(case.b) if B is None else B
    c = case.collapse
    n_iter = case.n // c
    idx = O.rotation_indices(case.a[:Bn], case.log_q, case.N, c, case.n_lut)
    s = np.asarray(case.s[: n_iter * c]).reshape(n_iter, c)
    p = (s * (1 << (c - 1 - np.arange(c)))).sum(axis=1)
    e = idx[:, np.arange(n_iter), p].astype(np.int64).sum(axis=1)
    bms = O.body_rotation(case.b[:Bn], case.log_q, case.N, case.n_lut).reshape(-1)
    return np.mod(e - bms, 2 * case.N)


def lut_at_switched_phase(case: PBSCase, B: int | None = None) -> np.ndarray:
    """(B, n_lut) messages the bootstrap outputs carry without noise:
    coefficients 0..n_lut-1 of X^e * T (T = the test polynomial's body row),
    over Delta_out, mod 2^(msg_bits+1).  The output noise (the key's) is far
    below Delta_out/2 in these cases, so decryptions must equal these exactly;
    whether they also equal LUT(m) depends on the modulus switch alone."""
    e = switched_rotation(case, B)
    T = np.asarray(case.lut[1], np.int64)
    vals = np.stack([O.coef_rotation(T, int(x))[: case.n_lut] for x in e])
    mod = 1 << (case.msg_bits + 1)
    return np.mod(np.rint(vals / case.delta_out).astype(np.int64), mod)
