"""LWE scheme of the reference's schemas/LWE_jax.py.

Key generation and encryption run on the host (NumPy, own seeded
generator: `key` is a numpy Generator or an integer seed where the reference
takes a JAX PRNG key).  key_switch_LWE runs on the GPU through the C ABI.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from cuTFHE import _core
from schemas.polynomial_jax import GLEV_polynomial, centered_mod


def rng_of(key) -> np.random.Generator:
    return key if isinstance(key, np.random.Generator) else np.random.default_rng(key)


def log2_exact(v: float) -> int:
    return int(math.log2(v) + 1e-6)


def uniform_torus(rng: np.random.Generator, shape, log_q: int) -> np.ndarray:
    """jax.random.randint(key, shape, -q/2, q/2) of the reference, as int64."""
    if log_q >= 64:
        return rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, size=shape,
                            dtype=np.int64, endpoint=True)
    half = 1 << (log_q - 1)
    return rng.integers(-half, half, size=shape, dtype=np.int64)


def encrypt_LWE(m, sk, dict_params: dict, key):
    """LWE_jax.py:9-28: a uniform, e = round(N(0,1))*sigma (the reference's
    own noise rule), b = centred_mod(<s,a> + m + e, q).  Batched over the
    leading axes of m.  Returns float64 (a (..., n), b (..., 1))."""
    rng = rng_of(key)
    q = dict_params["q"]
    log_q = log2_exact(q)
    sigma = dict_params["sigma"]
    m = np.asarray(m, dtype=np.float64)
    sk = np.asarray(sk, dtype=np.int64)
    a = uniform_torus(rng, m.shape + (len(sk),), log_q)
    e = np.rint(rng.standard_normal(m.shape)) * sigma
    dot = (a @ sk).astype(np.float64)          # exact int64 sum (wraps mod 2^64)
    b = centered_mod(dot + m + e, q)
    return a.astype(np.float64), b[..., None]


def decrypt_LWE(ciphertext, sk, dict_params: dict):
    """LWE_jax.py:30-48: round((b - <a,s>)/delta) centred mod t."""
    t = dict_params["t"]
    delta = dict_params["q"] / t
    a, b = np.asarray(ciphertext[0], np.float64), np.asarray(ciphertext[1], np.float64)
    phase = b.reshape(a.shape[:-1]) - a @ np.asarray(sk, np.float64)
    return centered_mod(np.rint(phase / delta), t)


def decrypt_LWE_quantization(ciphertext, sk, dict_params: dict, scale: float | None = None):
    """LWE_jax.py:52-62: centred phase divided by scale."""
    q = dict_params["q"]
    scale = q / dict_params["t"] if scale is None else scale
    a, b = np.asarray(ciphertext[0], np.float64), np.asarray(ciphertext[1], np.float64)
    phase = b.reshape(a.shape[:-1]) - a @ np.asarray(sk, np.float64)
    return centered_mod(phase, q) / scale


def get_keyswitch_sk_LWE(new_sk, old_sk, dict_params: dict, key):
    """LWE_jax.py:86-93: KSK[j,i] = LWE_new(-old_sk[i] * q/beta^(j+1)).

    Returns float64 (pk (l, N_in, n), b (l, N_in, 1))."""
    beta = dict_params["beta_ks"]
    l = dict_params["l_ks"]
    q = dict_params["q"]
    glev = GLEV_polynomial(np.asarray(old_sk, np.float64), beta, l, q)    # (l, N_in)
    return encrypt_LWE(-glev, new_sk, dict_params, key)


def key_switch_LWE(key_switching_key, ciphertext, dict_params: dict):
    """LWE_jax.py:97-117 on the GPU, batched over leading axes (the reference
    vmaps it, nn_tfhe/activation.py:150).

    key_switching_key (pk (l, N_in, n), b (l, N_in, 1)) float64 CUDA tensors;
    ciphertext (a (..., N_in), b (..., 1)).  Returns (a' (..., n), b' (..., 1))."""
    beta = dict_params["beta_ks"]
    l = dict_params["l_ks"]
    q = dict_params["q"]
    ksk_a, ksk_b = key_switching_key
    a, b = ciphertext
    _core.require_device(ksk_a, ksk_b, a, b)
    n_in = a.shape[-1]
    n_out = ksk_a.shape[-1]
    if tuple(ksk_a.shape) != (l, n_in, n_out):
        raise _core.TFHEInvalidArgument(f"KSK shape {tuple(ksk_a.shape)} != {(l, n_in, n_out)}")
    lead = a.shape[:-1]
    a2 = a.reshape(-1, n_in).to(torch.float64).contiguous()
    b2 = b.reshape(-1).to(torch.float64).contiguous()
    if b2.numel() != a2.shape[0]:
        raise _core.TFHEInvalidArgument("one body per mask expected")
    ka = ksk_a.to(torch.float64).contiguous()
    kb = ksk_b.reshape(l, n_in).to(torch.float64).contiguous()
    a_out = torch.empty(a2.shape[0], n_out, dtype=torch.float64, device=a.device)
    b_out = torch.empty(a2.shape[0], dtype=torch.float64, device=a.device)
    _core.check(_core.lib().bslstm_key_switch_lwe(
        _core.stream_of(a2), a2.data_ptr(), b2.data_ptr(), ka.data_ptr(), kb.data_ptr(),
        a2.shape[0], n_in, n_out, log2_exact(beta), l, log2_exact(q), a_out.data_ptr(),
        b_out.data_ptr()))
    return a_out.reshape(*lead, n_out), b_out.reshape(*lead, 1)
