"""RLWE text operations of the reference's schemas/text_jax.py that sit on the
bootstrapping path, as torch operations on the GPU.  (The hot many-LUT tail
of the bootstrap uses the fused HIP kernel in schemas/RLWE_jax.py instead.)
"""
from __future__ import annotations

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


def rotate_ciphertext(ciphertext, exponent):
    """text_jax.py:348-362: multiply both rows by X^exponent (one exponent per
    leading row)."""
    A, B = ciphertext[0], ciphertext[1]
    lead = A.shape[:-1]
    e = torch.as_tensor(exponent, device=A.device).reshape(-1)
    rows = A.reshape(-1, A.shape[-1]).shape[0]
    e = e.expand(rows) if e.numel() == 1 else e
    return (coef_rotation(A, e).reshape(*lead, -1), coef_rotation(B, e).reshape(*lead, -1))
