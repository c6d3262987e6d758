"""RLWE text operations of the reference's schemas/text_jax.py that sit on the
bootstrapping path, as torch operations on the GPU.  (The hot many-LUT tail
of the bootstrap uses the fused HIP kernel in schemas/RLWE_jax.py instead.)
"""
from __future__ import annotations

import math

import torch

from cuTFHE.utils import all_binary_vectors  # noqa: F401  (text_jax.py:440-454 twin)
from schemas.polynomial_jax import coef_rotation


def sample_extract(ciphertext, alpha: int):
    """text_jax.py:191-208: LWE of coefficient alpha of an RLWE (A, B),
    batched over leading axes.  a'_j = A[alpha-j] (j <= alpha), else
    -A[alpha-j+N]; b' = B[alpha] (shape (..., 1))."""
    A, B = ciphertext[0], ciphertext[1]
    N = A.shape[-1]
    j = torch.arange(N, device=A.device)
    src = torch.where(j <= alpha, alpha - j, alpha - j + N)
    sign = torch.where(j <= alpha, 1, -1).to(A.dtype)
    return A[..., src] * sign, B[..., alpha:alpha + 1]


def rotate_ciphertext(ciphertext, exponent, log_q: int | None = None):
    """text_jax.py:348-362: multiply both rows by X^exponent (one exponent per
    leading row).  int64 ciphertexts go through the HIP rotation kernel
    (log_q defaults to 64: the result is the same class mod any q)."""
    A, B = ciphertext[0], ciphertext[1]
    if A.dtype == torch.int64:
        from cuTFHE import torus
        lq = 64 if log_q is None else log_q
        return torus.rotate_poly(A, exponent, lq), torus.rotate_poly(B, exponent, lq)
    lead = A.shape[:-1]
    e = torch.as_tensor(exponent, device=A.device).reshape(-1)
    rows = A.reshape(-1, A.shape[-1]).shape[0]
    e = e.expand(rows) if e.numel() == 1 else e
    return (coef_rotation(A, e).reshape(*lead, -1), coef_rotation(B, e).reshape(*lead, -1))


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
