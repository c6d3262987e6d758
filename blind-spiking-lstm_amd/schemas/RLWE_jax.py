"""RLWE scheme and the programmable-bootstrap entry points of the reference's
schemas/RLWE_jax.py.

Host side (NumPy, own seeded generator in place of JAX PRNG keys): secret
keys, RLWE encryption/decryption and the collapsed bootstrapping key.
GPU side (C ABI, hand-written HIP): cuboot_merge, BR and bootstrapping.
"""
from __future__ import annotations

import numpy as np
import torch

from cuTFHE import _core
from cuTFHE.multiply import mul_bootstrap, prepare_bsk, rotation_indices, log2_int
from cuTFHE.utils import all_binary_vectors
from schemas.LWE_jax import EXACT_FLOAT_LOG_Q, log2_exact, rng_of, uniform_torus
from schemas.polynomial_jax import (GLEV_polynomial, apply_automorphism, centered_mod,
                                    float_to_torus, glev_messages, jax_fourier,
                                    negacyclic_mul_small, torus_mod)
from schemas.text_jax import rotate_ciphertext


def get_delta(dict_params: dict) -> float:
    """RLWE_jax.py:17-20."""
    return float(np.rint(dict_params["q"] / dict_params["t"]))


def sample_sk(dict_params: dict, key=None) -> np.ndarray:
    """Binary secret of length degree (RLWE_jax.py:24-29)."""
    return rng_of(key).choice([0.0, 1.0], size=(dict_params["degree"],))


def encrypt(m, sk, dict_params: dict, key):
    """RLWE_jax.py:36-54: A uniform, b = centred_mod(A*S + m + round(N(0,1)*sigma), q).

    Batched over leading axes of m; the product A*S is exact.  Returns
    float64 (A, b)."""
    rng = rng_of(key)
    q = dict_params["q"]
    log_q = log2_exact(q)
    sigma = dict_params["sigma"]
    if log_q > EXACT_FLOAT_LOG_Q:
        # int64 classes: float64 cannot hold them (the reference's float64
        # carriers round them); m as int64 or integer-valued float
        m = np.asarray(m)
        mi = m.astype(np.int64) if m.dtype.kind in "iu" else float_to_torus(m, log_q)
        A = uniform_torus(rng, mi.shape, log_q)
        e = np.rint(rng.standard_normal(mi.shape) * sigma).astype(np.int64)
        with np.errstate(over="ignore"):
            b = torus_mod(negacyclic_mul_small(A, sk, log_q) + mi + e, log_q)
        return A, b
    m = np.asarray(m, np.float64)
    A = uniform_torus(rng, m.shape, log_q)
    e = np.rint(rng.standard_normal(m.shape) * sigma)
    prod = negacyclic_mul_small(A, sk, log_q).astype(np.float64)
    b = centered_mod(prod + m + e, q)
    return A.astype(np.float64), b


def _phase(ciphertext, sk, dict_params):
    q = dict_params["q"]
    A, b = ciphertext[0], ciphertext[1]
    if isinstance(A, torch.Tensor):
        A, b = A.cpu().numpy(), b.cpu().numpy()
    A, b = np.asarray(A), np.asarray(b)
    if A.dtype == np.int64:
        log_q = log2_exact(q)
        with np.errstate(over="ignore"):
            ph = b.astype(np.int64) - negacyclic_mul_small(A, sk, log_q)
        return torus_mod(ph, log_q).astype(np.float64)
    b = np.asarray(b, np.float64)
    prod = negacyclic_mul_small(np.rint(A).astype(np.int64), sk, log2_exact(q)).astype(np.float64)
    return centered_mod(b - prod, q)


def decrypt(ciphertext, sk, dict_params: dict):
    """RLWE_jax.py:56-76."""
    t = dict_params["t"]
    delta = dict_params["q"] / t
    return centered_mod(np.rint(_phase(ciphertext, sk, dict_params) / delta), t)


def decrypt_quantization(ciphertext, sk, dict_params: dict, scale: float | None = None):
    """RLWE_jax.py:78-89."""
    scale = dict_params["q"] / dict_params["t"] if scale is None else scale
    return _phase(ciphertext, sk, dict_params) / scale


def get_boostrapping_key_collapse(sk_poly, sk_lwe, dict_params: dict, key, collapse: int):
    """Collapsed bootstrapping key (RLWE_jax.py:162-197).

    For group g and pattern p, ind = 1[sk_lwe[g*c:(g+1)*c] == p]; the left
    RGSW rows are RLWE(-S*ind*q/beta^i), the right rows RLWE(ind*q/beta^i) on
    the constant coefficient.  Returns ((left_pk, left_b), (right_pk, right_b)),
    each float64 (n_eq, l, N)."""
    rng = rng_of(key)
    q = dict_params["q"]
    beta = dict_params["beta_bs"]
    l = dict_params["l_bs"]
    sk_poly = np.asarray(sk_poly, np.float64)
    kron = all_binary_vectors(collapse)
    groups = np.asarray(sk_lwe).reshape(-1, collapse)
    all_eq = np.all(groups[:, None, :] == kron[None, :, :], axis=-1).astype(np.float64).ravel()
    log_q = log2_exact(q)
    if log_q > EXACT_FLOAT_LOG_Q:
        lb = log2_exact(beta)
        S = np.asarray(sk_poly, np.int64)
        ind = all_eq.astype(np.int64)
        left_msg = np.transpose(glev_messages(np.outer(ind, -S), lb, l, log_q), (1, 0, 2))
        right = np.zeros((len(ind), S.shape[-1]), np.int64)
        right[:, 0] = ind
        right_msg = np.transpose(glev_messages(right, lb, l, log_q), (1, 0, 2))
        return (encrypt(left_msg, sk_poly, dict_params, rng),
                encrypt(right_msg, sk_poly, dict_params, rng))
    # GLEV_polynomial per indicator row: (n_eq, l, N)
    left_msg = centered_mod(np.stack([GLEV_polynomial(row, beta, l, q)
                                      for row in np.outer(all_eq, -sk_poly)]), q)
    right = np.zeros((len(all_eq), sk_poly.shape[-1]))
    right[:, 0] = all_eq
    right_msg = centered_mod(np.stack([GLEV_polynomial(row, beta, l, q) for row in right]), q)
    left = encrypt(left_msg, sk_poly, dict_params, rng)
    right_ct = encrypt(right_msg, sk_poly, dict_params, rng)
    return left, right_ct


def get_keyswitch_sk(new_sk, old_sk, dict_params: dict, key):
    """RLWE_jax.py:91-98: GLev under new_sk of -old_sk * q/beta^i (i = 1..l):
    returns (A (l, N), b (l, N))."""
    rng = rng_of(key)
    q = dict_params["q"]
    log_q = log2_exact(q)
    l = dict_params["l_ks"]
    lb = log2_exact(dict_params["beta_ks"])
    msg = glev_messages(-np.asarray(old_sk, np.int64), lb, l, log_q)
    return encrypt(msg if log_q > EXACT_FLOAT_LOG_Q else msg.astype(np.float64), new_sk,
                   dict_params, rng)


def get_galois_key(sk, dict_params: dict, key) -> np.ndarray:
    """RLWE_jax.py:217-226: key i switches S(X^k_i) -> S, k_i = 2^(log2 N - i)+1,
    stacked to (log2 N, l, 2, N) and transformed: complex128 (log2 N, l, 2, N/2)."""
    rng = rng_of(key)
    N = dict_params["degree"]
    logn = int(np.log2(N) + 1e-6)
    keys = []
    for i in range(logn):
        sk_auto = apply_automorphism(np.asarray(sk, np.int64), 2 ** (logn - i) + 1)
        A, b = get_keyswitch_sk(sk, sk_auto, dict_params, rng)
        keys.append(np.stack([np.asarray(A), np.asarray(b)], axis=1))
    return jax_fourier(np.stack(keys).astype(np.float64))


def get_packing_KSK(lwe_sk, rlwe_sk, dict_params: dict, q, beta, l: int, key):
    """RLWE_jax.py:205-214: KSK[j, i] = RLWE_{rlwe_sk}(lwe_sk[j] * q/beta^(i+1)
    on the constant coefficient); returns (A (n, l, N), b (n, l, N))."""
    rng = rng_of(key)
    log_q = log2_exact(q)
    n = len(lwe_sk)
    N = len(rlwe_sk)
    poly = np.zeros((n, N), np.int64)
    poly[:, 0] = np.asarray(lwe_sk, np.int64)
    msg = np.transpose(glev_messages(poly, log2_exact(beta), l, log_q), (1, 0, 2))
    return encrypt(msg if log_q > EXACT_FLOAT_LOG_Q else msg.astype(np.float64), rlwe_sk,
                   dict_params, rng)


def bsk_to_fourier(BSK) -> np.ndarray:
    """nn_tfhe/activation.py:49-53: FFT of jnp.array(BSK) (2,2,n_eq,l,N), then
    permute to (2[pk/b], 2[L/R], l, n_eq, N/2) complex128."""
    arr = np.asarray([[np.asarray(c) for c in side] for side in BSK], dtype=np.float64)
    return np.transpose(jax_fourier(arr), (1, 0, 3, 2, 4))


def _stack_lut(encrypted_LUT) -> torch.Tensor:
    if isinstance(encrypted_LUT, (list, tuple)):
        return torch.stack(tuple(encrypted_LUT), dim=-2)
    return encrypted_LUT


def _blind_rotate_int(pk, lut_ct, bsk, q, beta, l, degree, collapse, all_rot, n_lut):
    log_q = log2_int(q)
    log_beta = log2_int(beta)
    if pk.dtype == torch.int64:
        from cuTFHE import torus
        idx = torus.rotation_indices(pk, log_q, degree, collapse, n_lut)
    else:
        idx = rotation_indices(pk, log_q, degree, collapse, n_lut)
    bsk_k = prepare_bsk(bsk, pk.shape[-1], collapse)
    return mul_bootstrap(lut_ct.to(torch.int64), bsk_k, all_rot, idx, log_beta, log_q, l,
                         form="auto")


def rotate_extract(br_out: torch.Tensor, b: torch.Tensor, log_q: int, n_lut: int):
    """Fused X^(-b_ms) rotation + sample extraction of coefficients
    0..n_lut-1 (RLWE_jax.py:295-303) -> (masks (B,n_lut,N), bodies (B,n_lut))."""
    _core.require_device(br_out, b)
    if b.dtype == torch.int64:
        from cuTFHE import torus
        return torus.rotate_extract(br_out, b, log_q, n_lut)
    B, _, N = br_out.shape
    br = br_out.contiguous()
    body = b.reshape(-1).to(torch.float64).contiguous()
    if body.numel() != B:
        raise _core.TFHEInvalidArgument("one LWE body per blind-rotated ciphertext expected")
    masks = torch.empty(B, n_lut, N, dtype=torch.float64, device=br.device)
    bodies = torch.empty(B, n_lut, dtype=torch.float64, device=br.device)
    lib = _core.lib()
    chunk = 65535
    for s in range(0, B, chunk):
        e = min(B, s + chunk)
        _core.check(lib.bslstm_rotate_extract(
            _core.stream_of(br), br[s:e].data_ptr(), body[s:e].data_ptr(), e - s, N, log_q, n_lut,
            masks[s:e].data_ptr(), bodies[s:e].data_ptr()))
    return masks, bodies


def cuboot_merge(c_lwe_ks, encrypted_LUT, bsk, q, beta, l, degree, collapse,
                 all_rot_possible_fourier, n_lut):
    """Programmable bootstrap of a batch of LWE ciphertexts
    (RLWE_jax.py:288-303): blind-rotate each ciphertext's test polynomial,
    rotate by -b_ms = -rint(2N*b/(q*n_lut))*n_lut and extract coefficients
    0..n_lut-1.

    c_lwe_ks = (pk (B, n), b (B, 1)); encrypted_LUT = [(B, N), (B, N)].
    Returns (masks (B, n_lut, N), bodies (B, n_lut, 1)) float64, or
    ((B, N), (B, 1)) when n_lut is None."""
    pk, b = c_lwe_ks
    lut = _stack_lut(encrypted_LUT)
    if pk.dtype == torch.int64 and lut.dtype != torch.int64:
        from cuTFHE.torus import _i64
        lut = _i64(lut)
    nl = 1 if n_lut is None else n_lut
    br = _blind_rotate_int(pk, lut, bsk, q, beta, l, degree, collapse,
                           all_rot_possible_fourier, nl)
    masks, bodies = rotate_extract(br, b, log2_int(q), nl)
    if n_lut is None:
        return masks[:, 0], bodies[:, 0:1]
    return masks, bodies[..., None]


def BR(c_lwe_ks, encrypted_LUT, bsk, q, beta, l, degree, collapse, all_rot_possible_fourier,
       n_lut):
    """RLWE_jax.py:305-315: the rotated test polynomials (B, 2, N) before
    extraction."""
    pk, b = c_lwe_ks
    lut = _stack_lut(encrypted_LUT)
    br = _blind_rotate_int(pk, lut, bsk, q, beta, l, degree, collapse, all_rot_possible_fourier,
                           n_lut).to(torch.float64)
    b_ms = torch.round((2 * degree) * b.reshape(-1).to(torch.float64) / (q * n_lut)) * n_lut
    A, Bp = rotate_ciphertext((br[:, 0], br[:, 1]), -b_ms)
    return torch.stack([A, Bp], dim=1)


def bootstrapping(lwe_ciphertext, LUT, BSK, q, beta, l, degree, collapse, all_rot_fft=None,
                  n_lut: int = 1):
    """RLWE_jax.py:234-274 (the reference's pure-JAX path) on the GPU path:
    same mathematics, rotation applied after the blind rotation."""
    pk, b = lwe_ciphertext
    single = pk.dim() == 1
    if single:
        pk, b = pk[None], b.reshape(1, 1)
        LUT = [x[None] for x in LUT] if isinstance(LUT, (list, tuple)) else LUT[None]
    masks, bodies = cuboot_merge((pk, b), LUT, BSK, q, beta, l, degree, collapse, all_rot_fft,
                                 None if n_lut == 1 else n_lut)
    if single:
        return masks[0], bodies[0]
    return masks, bodies
