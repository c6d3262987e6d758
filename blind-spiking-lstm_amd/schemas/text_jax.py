"""RLWE text operations of the reference's schemas/text_jax.py that sit on the
bootstrapping path, on the GPU through the HIP kernels of cuTFHE.torus.
float64 tensors (the reference's carrier of torus values, integer-valued and
exact below 2^53) go through the int64 kernels and come back as float64.
(The hot many-LUT tail of the bootstrap uses the fused HIP kernel in
schemas/RLWE_jax.py.)
"""
from __future__ import annotations

import math

import torch

from cuTFHE.utils import all_binary_vectors  # noqa: F401  (text_jax.py:440-454 twin)


def _as_carrier(x: torch.Tensor, like: torch.Tensor) -> torch.Tensor:
    """int64 kernel result back in the caller's dtype (float64 stays float64)."""
    return x if like.dtype == torch.int64 else x.to(like.dtype)


def sample_extract(ciphertext, alpha: int):
    """text_jax.py:191-208: LWE of coefficient alpha of an RLWE (A, B),
    batched over leading axes.  a'_j = A[alpha-j] (j <= alpha), else
    -A[alpha-j+N]; b' = B[alpha] (shape (..., 1)).  Runs as X^(-alpha)
    rotation + extraction of coefficient 0 (the same LWE) on the HIP
    rotate/extract kernels; no reduction mod q (log q = 64 wraps nothing
    below 2^63)."""
    from cuTFHE import torus
    A, B = ciphertext[0], ciphertext[1]
    lead, N = A.shape[:-1], A.shape[-1]
    ct = torch.stack([torus._i64(A).reshape(-1, N), torus._i64(B).reshape(-1, N)], dim=1)
    ct = torus.rotate_poly(ct, -int(alpha), 64)
    masks, bodies = torus.rotate_extract(ct, None, 64, 1)
    return (_as_carrier(masks[:, 0].reshape(*lead, N), A),
            _as_carrier(bodies.reshape(*lead, 1), B))


def rotate_ciphertext(ciphertext, exponent, log_q: int | None = None):
    """text_jax.py:348-362: multiply both rows by X^exponent (one exponent per
    leading row) on the HIP rotation kernel.  log_q defaults to 64: int64
    classes are the same mod any q, and float64 values (|x| < 2^53) come back
    unreduced, as the reference's float rotation leaves them."""
    from cuTFHE import torus
    A, B = ciphertext[0], ciphertext[1]
    lq = 64 if log_q is None else log_q
    e = exponent
    if isinstance(e, torch.Tensor) and e.dtype != torch.int64:
        e = torch.round(e.to(torch.float64)).to(torch.int64)
    return (_as_carrier(torus.rotate_poly(A, e, lq), A), _as_carrier(torus.rotate_poly(B, e, lq), B))


# ---------------------------------------------------------------------------
# ciphertext arithmetic and packing on int64 torus values (HIP kernels)
# ---------------------------------------------------------------------------


def _log_q(q) -> int:
    return int(math.log2(q) + 1e-6)


def sum_ciphertext_ciphertext(ciphertext1, ciphertext2, q, modulo: bool = True):
    """text_jax.py:27-43.  int64 tensors: exact, always reduced (the
    reference's modulo=False keeps float values above q/2; the class mod q is
    the same)."""
    from cuTFHE import torus
    lq = _log_q(q)
    return (torus.axpy(ciphertext1[0], ciphertext2[0], 1, 1, lq),
            torus.axpy(ciphertext1[1], ciphertext2[1], 1, 1, lq))


def centered_mod_ciphertext(ciphertext, q):
    """text_jax.py:45-57."""
    from cuTFHE import torus
    lq = _log_q(q)
    return tuple(torus.axpy(c, torch.zeros_like(c), 1, 0, lq) for c in ciphertext[:2])


def sum_plaintext_to_ciphertext(plaintext, ciphertext):
    """text_jax.py:12-24: (A, B + m)."""
    return ciphertext[0], ciphertext[1] + plaintext


def multiply_plaintext_ciphertext(plaintext, ciphertext, degree: int, q=2.0 ** 64):
    """text_jax.py:59-78 (fft_product_twice, rounded) for integer plaintext
    polynomials, batched over leading axes: (P(X)*A, P(X)*B)."""
    from cuTFHE import torus
    A, B = ciphertext[0], ciphertext[1]
    N = A.shape[-1]
    lead = A.shape[:-1]
    ct = torch.stack([torus._i64(A).reshape(-1, N), torus._i64(B).reshape(-1, N)], dim=1)
    P = torus._i64(torch.as_tensor(plaintext, device=A.device)).reshape(-1, N)
    P = P.expand(ct.shape[0], N) if P.shape[0] == 1 else P
    out = torus.lut_product_sum(ct.reshape(1, 1, *ct.shape), torus.poly_fft(P.contiguous()), 1,
                                _log_q(q))[0]
    return out[:, 0].reshape(*lead, N), out[:, 1].reshape(*lead, N)


def pack_lwe_to_rlwe(lwe, galois_keys, beta, l: int, q, degree: int):
    """text_jax.py:314-345: trace packing of LWEs (a (..., N), b (..., 1)) into
    RLWEs (A (..., N), B (..., N)); galois_keys complex128 (log2 N, l, 2, N/2)
    on the GPU."""
    from cuTFHE import torus
    a, b = lwe
    out = torus.pack_lwe_to_rlwe(a, b, galois_keys, int(math.log2(beta) + 1e-6), l, _log_q(q))
    return out[..., 0, :], out[..., 1, :]


def prepare_ksp_fft(KSP) -> torch.Tensor:
    """Transform a packing key (A (n, l, N), b (n, l, N)) (get_packing_KSK)
    into the (n, l, 2, N/2) complex128 layout the packing kernel reads."""
    from cuTFHE import torus
    if isinstance(KSP, torch.Tensor) and KSP.dtype == torch.complex128:
        return KSP.contiguous()
    A, B = KSP
    st = torch.stack([torus._i64(A), torus._i64(B)], dim=2).contiguous()    # (n, l, 2, N)
    return torus.poly_fft(st)


def packing(ciphertexts, KSP, q, beta, l: int, degree: int):
    """text_jax.py:211-239: pack LWEs (masks (..., n_lwes, n), bodies
    (..., n_lwes, 1)) into one RLWE each (A (..., N), B (..., N)); LWE i lands
    on coefficient i.  KSP: the packing key, as returned by get_packing_KSK
    or already through prepare_ksp_fft."""
    from cuTFHE import torus
    masks, bodies = ciphertexts[0], ciphertexts[1]
    lead = masks.shape[:-2]
    n_lwes, n_in = masks.shape[-2], masks.shape[-1]
    kf = prepare_ksp_fft(KSP)
    out = torus.packing(masks.reshape(-1, n_lwes, n_in), bodies.reshape(-1, n_lwes), kf,
                        int(math.log2(beta) + 1e-6), l, _log_q(q))
    return out[:, 0].reshape(*lead, degree), out[:, 1].reshape(*lead, degree)
