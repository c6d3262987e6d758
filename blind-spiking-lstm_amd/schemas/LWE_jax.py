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
from schemas.polynomial_jax import (GLEV_polynomial, centered_mod, float_to_torus, glev_messages,
                                    torus_mod)

# above this modulus float64 cannot hold torus classes: int64 carriers
EXACT_FLOAT_LOG_Q = 53


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
    sk = np.asarray(sk, dtype=np.int64)
    if log_q > EXACT_FLOAT_LOG_Q:
        # int64 classes (see EXACT_FLOAT_LOG_Q): m as int64 or integer-valued float
        m = np.asarray(m)
        mi = m.astype(np.int64) if m.dtype.kind in "iu" else float_to_torus(m, log_q)
        a = uniform_torus(rng, mi.shape + (len(sk),), log_q)
        e = (np.rint(rng.standard_normal(mi.shape)) * sigma).astype(np.int64)
        with np.errstate(over="ignore"):
            b = torus_mod((a * sk).sum(axis=-1) + mi + e, log_q)
        return a, b[..., None]
    m = np.asarray(m, dtype=np.float64)
    a = uniform_torus(rng, m.shape + (len(sk),), log_q)
    e = np.rint(rng.standard_normal(m.shape)) * sigma
    dot = (a @ sk).astype(np.float64)          # exact int64 sum (wraps mod 2^64)
    b = centered_mod(dot + m + e, q)
    return a.astype(np.float64), b[..., None]


def decrypt_LWE(ciphertext, sk, dict_params: dict):
    """LWE_jax.py:30-48: round((b - <a,s>)/delta) centred mod t."""
    t = dict_params["t"]
    delta = dict_params["q"] / t
    return centered_mod(np.rint(lwe_phase(ciphertext, sk, dict_params) / delta), t)


def lwe_phase(ciphertext, sk, dict_params: dict) -> np.ndarray:
    """Centred phase b - <a, s> mod q (float64 values; exact int64 arithmetic
    for int64 ciphertexts)."""
    a, b = ciphertext[0], ciphertext[1]
    if isinstance(a, torch.Tensor):
        a, b = a.cpu().numpy(), b.cpu().numpy()
    a, b = np.asarray(a), np.asarray(b)
    log_q = log2_exact(dict_params["q"])
    if a.dtype == np.int64:
        with np.errstate(over="ignore"):
            ph = b.reshape(a.shape[:-1]).astype(np.int64) - (a * np.asarray(sk, np.int64)).sum(-1)
        return torus_mod(ph, log_q).astype(np.float64)
    phase = b.reshape(a.shape[:-1]).astype(np.float64) - a.astype(np.float64) @ np.asarray(sk, np.float64)
    return centered_mod(phase, dict_params["q"])


def decrypt_LWE_quantization(ciphertext, sk, dict_params: dict, scale: float | None = None):
    """LWE_jax.py:52-62: centred phase divided by scale."""
    q = dict_params["q"]
    scale = q / dict_params["t"] if scale is None else scale
    return lwe_phase(ciphertext, sk, dict_params) / scale


def get_keyswitch_sk_LWE(new_sk, old_sk, dict_params: dict, key):
    """LWE_jax.py:86-93: KSK[j,i] = LWE_new(-old_sk[i] * q/beta^(j+1)).

    Returns float64 (pk (l, N_in, n), b (l, N_in, 1))."""
    beta = dict_params["beta_ks"]
    l = dict_params["l_ks"]
    q = dict_params["q"]
    if log2_exact(q) > EXACT_FLOAT_LOG_Q:
        glev = glev_messages(-np.asarray(old_sk, np.int64), log2_exact(beta), l, log2_exact(q))
        return encrypt_LWE(glev, new_sk, dict_params, key)
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
    ksk_a, ksk_b = key_switching_key[0], key_switching_key[1]
    tiles = key_switching_key[2] if len(key_switching_key) > 2 else None
    a, b = ciphertext
    _core.require_device(ksk_a, ksk_b, a, b)
    n_in = a.shape[-1]
    n_out = ksk_a.shape[-1]
    if ksk_a.dtype == torch.int64 or a.dtype == torch.int64 or log2_exact(q) > 32:
        # exact int64 forms (any log2 q <= 64); with the key's int8 limb tiles
        # (a third entry, torus.prepare_ks_key) on the matrix cores
        from cuTFHE import torus
        lead = a.shape[:-1]
        if tiles is not None:
            a_out, b_out = torus.key_switch_mfma(tiles, a.reshape(-1, n_in), b.reshape(-1), n_out,
                                                 log2_exact(beta), l, log2_exact(q))
        else:
            a_out, b_out = torus.key_switch(ksk_a, ksk_b, a.reshape(-1, n_in), b.reshape(-1),
                                            log2_exact(beta), l, log2_exact(q))
        return a_out.reshape(*lead, n_out), b_out.reshape(*lead, 1)
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
