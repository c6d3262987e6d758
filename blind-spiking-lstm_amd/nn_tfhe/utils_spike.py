"""Test-polynomial (LUT) builders of the reference's nn_tfhe/utils_spike.py.

Host-side NumPy; a test polynomial is built once per layer and moved to the
GPU with the keys.
"""
from __future__ import annotations

import numpy as np


def _rotate_np(poly: np.ndarray, alpha: int) -> np.ndarray:
    """poly * X^alpha mod X^N+1 (schemas/polynomial_jax.py:150-169)."""
    n = poly.shape[-1]
    rot = np.roll(poly, alpha, axis=-1) * (-1.0) ** ((alpha // n) % 2)
    return np.where(np.arange(n) < alpha % n, -rot, rot)


def get_plain_LUT(degree_lut: int, n_lut: int, index: int, s: float, gamma: float) -> np.ndarray:
    """utils_spike.py:72-87: sub-LUT `index` of the step 1[x >= gamma] on
    linspace(-s, s, N/n_lut), interleaved with stride n_lut and rotated by
    X^(-N/2) so the signed phase psi in [-N/2, N/2) reads entry psi + N/2."""
    sub = degree_lut // n_lut
    values = np.where(np.linspace(-s, s, num=sub, endpoint=False) >= gamma, 1.0, 0.0)
    full = np.zeros(degree_lut)
    full[np.arange(sub) * n_lut + index] = values
    return _rotate_np(full, -degree_lut // 2)


def many_lut_polynomial(degree_lut: int, fns, msg_bits: int, delta_out: float) -> np.ndarray:
    """Test polynomial packing len(fns) LUTs of signed b-bit messages
    x in [-2^(b-1), 2^(b-1)) encoded as x*q/2^(b+1), in the interleaved layout
    of get_plain_LUT: sub-LUT k at coefficients base*n_lut + k, base reading
    psi = base*n_lut - N/2.  Output k of the bootstrap is delta_out*fns[k](x).
    The upper half box holds -f(-2^(b-1)), the negacyclic image of the lower."""
    n_lut = len(fns)
    sub = degree_lut // n_lut
    width = degree_lut >> msg_bits
    base = np.arange(sub)
    x = np.floor((base * n_lut - degree_lut // 2) / width + 0.5).astype(np.int64)
    top = 1 << (msg_bits - 1)
    full = np.zeros(degree_lut)
    for k, f in enumerate(fns):
        vals = np.array([-f(int(v) - 2 * top) if v >= top else f(int(v)) for v in x], np.float64)
        full[base * n_lut + k] = vals * delta_out
    return _rotate_np(full, -degree_lut // 2)
