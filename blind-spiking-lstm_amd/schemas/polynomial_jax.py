"""Polynomial helpers of the reference's schemas/polynomial_jax.py.

Host-side (NumPy) pieces serve key generation and encryption; torch pieces
run on the GPU.  The module name is kept so that import lines written for
the reference resolve unchanged.
"""
from __future__ import annotations

import numpy as np
import torch


def centered_mod(x, q):
    """x - q*rint(x/q) (polynomial_jax.py:7-9); numpy or torch input."""
    if isinstance(x, torch.Tensor):
        return x - q * torch.round(x / q)
    return x - q * np.rint(x / q)


def _twist(N: int, sign: float) -> np.ndarray:
    k = np.arange(N)
    return np.cos(np.pi * k / N) + sign * 1j * np.sin(np.pi * k / N)


def jax_fourier(p) -> np.ndarray:
    """Negacyclic transform, first N/2 bins of FFT(exp(-i pi k/N) p)
    (polynomial_jax.py:11-15)."""
    p = np.asarray(p, dtype=np.float64)
    N = p.shape[-1]
    return np.fft.fft(_twist(N, -1.0) * p, axis=-1)[..., : N // 2]


def jax_inversefourier(F) -> np.ndarray:
    """Inverse of jax_fourier with the Hermitian extension, rounded
    (polynomial_jax.py:17-21)."""
    F = np.asarray(F)
    full = np.concatenate([F, np.flip(np.conj(F), axis=-1)], axis=-1)
    n = full.shape[-1]
    return np.round(np.real(_twist(n, 1.0) * np.fft.ifft(full, axis=-1)))


def negacyclic_mul_small(a, s, log_q: int) -> np.ndarray:
    """Exact a(X)*s(X) mod (X^N+1, 2^log_q) for a key s with |s_i| <= 1.

    The role of the rounded FFT product the reference uses in encryption
    (fft_polynomial_multiply, polynomial_jax.py:26-43), made exact: a is cut
    into 16-bit limbs so every limb product stays below 2^(16+log2 N) and the
    FP64 transform rounds it exactly; limbs are recombined mod 2^64.
    """
    a = np.asarray(a, dtype=np.int64)
    s = np.asarray(s, dtype=np.float64)
    fs = jax_fourier(s)
    au = a.astype(np.uint64)
    acc = np.zeros(a.shape, dtype=np.uint64)
    for limb in range(4):
        part = ((au >> np.uint64(16 * limb)) & np.uint64(0xFFFF)).astype(np.float64)
        prod = jax_inversefourier(jax_fourier(part) * fs).astype(np.int64)
        acc += prod.astype(np.uint64) << np.uint64(16 * limb)
    q_mask = np.uint64((1 << log_q) - 1) if log_q < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    r = (acc & q_mask).astype(np.int64) if log_q < 64 else acc.astype(np.int64)
    if log_q < 64:
        half = 1 << (log_q - 1)
        r = np.where(r >= half, r - (1 << log_q), r)
    return r


def GLEV_polynomial(x, beta: float, l: int, q: float) -> np.ndarray:
    """x * q/beta^i for i = 1..l (polynomial_jax.py:115-134)."""
    basis = np.array([q / (beta ** i) for i in range(1, l + 1)])[:, None]
    return np.asarray(x, np.float64) * basis


def coef_rotation(polynomial: torch.Tensor, alpha) -> torch.Tensor:
    """polynomial * X^alpha mod X^N+1 (polynomial_jax.py:150-169), batched
    over the leading axis with one integer exponent per row (or one for all),
    on the HIP rotation kernel (bslstm_rotate_poly).  int64 or integer-valued
    float64 (exact below 2^53); the result keeps the input's dtype and is not
    reduced mod q, as the reference's rotation."""
    from cuTFHE import torus
    N = polynomial.shape[-1]
    if isinstance(alpha, torch.Tensor) and alpha.dtype != torch.int64:
        alpha = torch.round(alpha.to(torch.float64)).to(torch.int64)
    out = torus.rotate_poly(polynomial.reshape(-1, N), alpha, 64).reshape(polynomial.shape)
    return out if polynomial.dtype == torch.int64 else out.to(polynomial.dtype)


# ---------------------------------------------------------------------------
# integer torus helpers (moduli up to 2^64, where float64 cannot hold classes)
# ---------------------------------------------------------------------------


def torus_mod(x, log_q: int) -> np.ndarray:
    """Centred representative mod 2^log_q of int64 classes (integer form of
    centered_mod, polynomial_jax.py:7-9); sums wrap mod 2^64 beforehand."""
    x = np.asarray(x).astype(np.int64)
    s = 64 - log_q
    return ((x.astype(np.uint64) << np.uint64(s)).astype(np.int64) >> np.int64(s)).astype(np.int64)


def float_to_torus(v, log_q: int) -> np.ndarray:
    """int64 class mod 2^log_q of integer-valued float64 values, including
    magnitudes at or above 2^63 (e.g. s*q/beta with q = 2^64)."""
    v = np.asarray(v, np.float64)
    r = v - 2.0 ** 64 * np.rint(v / 2.0 ** 64)            # [-2^63, 2^63], exact
    r = np.where(r >= 2.0 ** 63, r - 2.0 ** 64, r)
    return torus_mod(r.astype(np.int64), log_q)


def glev_messages(x, log_beta: int, l: int, log_q: int) -> np.ndarray:
    """GLEV_polynomial (polynomial_jax.py:115-134) in exact integers:
    x * q/beta^i for i = 1..l, as int64 classes (l, *x.shape)."""
    x = np.asarray(x).astype(np.int64)
    if l * log_beta > log_q:
        raise ValueError("l*log2(beta) must not exceed log2(q)")
    return np.stack([torus_mod((x.astype(np.uint64) << np.uint64(log_q - i * log_beta)).astype(
        np.int64), log_q) for i in range(1, l + 1)])


def decomposition(gamma, beta, l: int, q) -> np.ndarray:
    """Float digit recursion of polynomial_jax.py:90-111 (host form):
    x = gamma/q; digit = round_half_even(x*beta); x = x*beta - digit."""
    x = np.asarray(gamma, np.float64) / q
    out = []
    for _ in range(l):
        xb = x * beta
        d = np.rint(xb)
        x = xb - d
        out.append(d)
    return np.stack(out)


def apply_automorphism(polynomial, alpha: int) -> np.ndarray:
    """p(X) -> p(X^alpha) in Z[X]/(X^N+1) (polynomial_jax.py:171-200)."""
    p = np.asarray(polynomial)
    N = p.shape[-1]
    e = (np.arange(N) * int(alpha)) % (2 * N)
    out = np.zeros_like(p)
    with np.errstate(over="ignore"):
        out[..., e % N] = np.where(e >= N, -p, p)
    return out


def weights_to_polynomial(weights) -> np.ndarray:
    """polynomial_jax.py:137-148: -flip(w), last entry replaced by w[0]."""
    w = np.asarray(weights).reshape(-1)
    p = -w[::-1].copy()
    p[-1] = w[0]
    return p
