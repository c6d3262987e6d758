"""Key and layer setup of the reference's setup_crypto.py.

Keys are drawn from a numpy Generator seeded with `seed` (the reference
mixes np.random with JAX PRNG keys, setup_crypto.py:17-37, which cannot be
reproduced without JAX).  The reference builds Linear layers and LUTs from a
trained plaintext model (setup_crypto.py:39-82); here the weights, scales and
thresholds are passed as arrays.
"""
from __future__ import annotations

import numpy as np

import parameters_crypto as P
from nn_tfhe.activation import HeavysideBoostrapper, ReluBoostrapper, prepare_bootstrapping
from nn_tfhe.Spike_LSTM import Linear
from nn_tfhe.utils_spike import get_plain_LUT
from schemas.RLWE_jax import get_galois_key, get_packing_KSK


def identity(x, degree):
    """setup_crypto.py:13-14."""
    return np.ones(degree)


def sample_crypto_keys(seed, params=None, beta_x=P.beta_x):
    """setup_crypto.py:17-37.  params = (dict_params, dict_params_lut,
    dict_params_ks_LWE, dict_params_packing) (default: parameters_crypto).
    Returns (sk, ksk_packing, galois_key, bsk, key_switching_key_bs, sk_bsk);
    the last (the small LWE key) is returned for decryption in tests."""
    dp, dpl, dpk, dpp = params if params is not None else (
        P.dict_params, P.dict_params_lut, P.dict_params_ks_LWE, P.dict_params_packing)
    rng = np.random.default_rng(seed)
    sk = rng.choice([0.0, 1.0], dp["degree"])
    sk_bsk = rng.choice([0.0, 1.0], dpk["degree"])
    sk_lut = sk
    ksk_packing = get_packing_KSK(sk_lut, sk_lut, dp, dp["q"], dp["beta_ks"], dp["l_ks"], rng)
    galois_key = get_galois_key(sk, dpp, rng)
    _, _, bsk, key_switching_key_bs, _ = prepare_bootstrapping(
        rng, sk, sk_lut, sk_bsk, dp, dpl, dpk, identity, beta_x, dpl["collapse"])
    return sk, ksk_packing, galois_key, bsk, key_switching_key_bs, sk_bsk


def make_linear(W, B, factor, dict_params, device="cuda") -> Linear:
    """Linear(hidden, dict_params, factor) + set_weights(W, B)
    (setup_crypto.py:39-67 for one layer)."""
    lin = Linear(np.asarray(W).shape[1], dict_params, factor, device=device)
    lin.set_weights(W, B)
    return lin


def get_lut(n_lut, hidden_dim, degree_lut, sf, si, so, gammaf, gammai, gammao, eta,
            beta_x=P.beta_x):
    """setup_crypto.py:69-82: lut_f, lut_o (n_lut, hidden/n_lut, N) plaintext
    step LUTs; lut_i = [zeros, round(sum_k lut_i_k * eta_k * beta_x)].
    sf/si/so: (hidden/n_lut,) scales; gamma*, eta: (n_lut, hidden/n_lut)."""
    G = hidden_dim // n_lut

    def build(s, gamma):
        return np.stack([np.stack([get_plain_LUT(degree_lut, n_lut, k, float(s[g]),
                                                 float(gamma[k][g])) for g in range(G)])
                         for k in range(n_lut)])

    lut_f = build(sf, gammaf)
    lut_i = build(si, gammai)
    lut_i = np.rint(np.sum(lut_i * np.asarray(eta, np.float64)[..., None], axis=0) * beta_x)
    lut_o = build(so, gammao)
    return lut_f, [np.zeros((G, degree_lut)), lut_i], lut_o


def get_tfhe_activation(bsk, key_switching_key_bs, ksk_packing, galois_key, s_head1, n_lut,
                        params=None, device="cuda"):
    """setup_crypto.py:86-118."""
    dp, dpl, dpk, dpp = params if params is not None else (
        P.dict_params, P.dict_params_lut, P.dict_params_ks_LWE, P.dict_params_packing)
    heavy = HeavysideBoostrapper(bsk, key_switching_key_bs, ksk_packing, galois_key, dp, dpl, dpk,
                                 dpp, dpl["collapse"], P.all_rot_possible_fourier, n_lut,
                                 device=device)
    relu = ReluBoostrapper(bsk, key_switching_key_bs, ksk_packing, dp, dpl, dpk, dpp,
                           dpl["collapse"], P.all_rot_possible_fourier, P.beta_x, s_head1,
                           device=device)
    return heavy, relu


# ---------------------------------------------------------------------------
# device key generation (serving: one key set per client, generated on the
# GPU in the layouts the kernels read)
# ---------------------------------------------------------------------------

def _wrap_const(v: int) -> int:
    """Python int -> the int64 with the same class mod 2^64."""
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= 1 << 63 else v


def _uniform64(shape, gen, device):
    import torch
    hi = torch.randint(-2 ** 31, 2 ** 31, shape, generator=gen, device=device, dtype=torch.int64)
    lo = torch.randint(0, 2 ** 32, shape, generator=gen, device=device, dtype=torch.int64)
    return hi * (1 << 32) + lo


def _negacyclic_mul_key(a, s):
    """Exact a(X)*s(X) mod (X^N+1, 2^64) for int64 a (..., N) and a binary
    key s (N,) on the device: four 16-bit limbs of a, each product through
    the engine's negacyclic transform pair (bslstm_poly_fft / _poly_ifft)
    and rounded (|coefficients| < 2^16 * N, exact)."""
    from cuTFHE import torus
    fs = torus.poly_fft(s.to(a.dtype))
    acc = None
    for limb in range(4):
        part = (a >> (16 * limb)) & 0xFFFF
        c = torus.poly_ifft(torus.poly_fft(part) * fs)
        term = c * _wrap_const(1 << (16 * limb))
        acc = term if acc is None else acc + term
    return acc


def _encrypt_rlwe_dev(msg, S, sigma, gen):
    import torch
    A = _uniform64(tuple(msg.shape), gen, msg.device)
    e = torch.round(torch.randn(tuple(msg.shape), generator=gen, device=msg.device,
                                dtype=torch.float64) * sigma).to(torch.int64)
    return A, _negacyclic_mul_key(A, S) + msg + e


def _glev_dev(x, log_beta: int, l: int):
    """x * 2^(64 - i*log_beta), i = 1..l, as int64 classes mod 2^64: (l, *x.shape)."""
    import torch
    return torch.stack([x * _wrap_const(1 << (64 - i * log_beta)) for i in range(1, l + 1)])


def sample_client_keys_device(seed: int, params=None, device="cuda"):
    """One client's keys at q = 2^64, generated on the device (own seeded
    torch generator).  Returns (keys, secrets): keys = dict with
      bsk    complex128 (n_iter, 2, l, 2, 2^c, N/2)   collapsed BSK, kernel layout
      ks     int8 tiles of the MFMA key switch (N -> n)
      gk     complex128 (log2 N, l, 2, N/2)           Galois keys (trace packing)
      ksp    complex128 (N, l, 2, N/2)                packing key
    and secrets = (sk (N,), sk_bsk (n,)) as numpy arrays (client side)."""
    import numpy as np
    import torch
    from cuTFHE import torus
    from cuTFHE.utils import all_binary_vectors
    dp, dpl, dpk, dpp = params if params is not None else (
        P.dict_params, P.dict_params_lut, P.dict_params_ks_LWE, P.dict_params_packing)
    if dp["q"] != 2.0 ** 64:
        raise ValueError("device key generation is written for q = 2^64")
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    N, n = dp["degree"], dpk["degree"]
    c = dpl["collapse"]
    S = torch.randint(0, 2, (N,), generator=gen, device=dev, dtype=torch.int64)
    s = torch.randint(0, 2, (n,), generator=gen, device=dev, dtype=torch.int64)
    # collapsed BSK (RLWE_jax.py:162-197)
    lb, L = int(np.log2(dpl["beta_bs"]) + 1e-6), dpl["l_bs"]
    kron = torch.as_tensor(all_binary_vectors(c), device=dev, dtype=torch.int64)   # (m, c)
    ind = (s.reshape(-1, 1, c) == kron[None]).all(-1).to(torch.int64)            # (n_iter, m)
    n_iter, m = ind.shape
    left = _glev_dev(-ind[..., None] * S, lb, L)                                  # (L, n_iter, m, N)
    right_poly = torch.zeros(n_iter, m, N, dtype=torch.int64, device=dev)
    right_poly[..., 0] = ind
    right = _glev_dev(right_poly, lb, L)
    coef = torch.empty(n_iter, 2, L, 2, m, N, dtype=torch.int64, device=dev)
    for side, msg in ((0, left), (1, right)):
        A, b = _encrypt_rlwe_dev(msg, S, dpl["sigma"], gen)
        coef[:, 0, :, side] = A.permute(1, 0, 2, 3)
        coef[:, 1, :, side] = b.permute(1, 0, 2, 3)
    bsk = torus.poly_fft(coef)
    del coef, left, right, right_poly
    # LWE key-switching key N -> n (LWE_jax.py:86-93)
    lbk, lk = int(np.log2(dpk["beta_ks"]) + 1e-6), dpk["l_ks"]
    msg = _glev_dev(-S, lbk, lk)                                                  # (lk, N)
    ka = _uniform64((lk, N, n), gen, dev)
    e = (torch.round(torch.randn((lk, N), generator=gen, device=dev, dtype=torch.float64)) *
         dpk["sigma"]).to(torch.int64)
    kb = (ka * s).sum(-1) + msg + e
    ks = torus.prepare_ks_key(ka, kb)
    del ka, kb
    # Galois keys (RLWE_jax.py:217-226)
    lbp, lp = int(np.log2(dpp["beta_ks"]) + 1e-6), dpp["l_ks"]
    logn = int(np.log2(N) + 1e-6)
    S_np = S.cpu().numpy()
    gks = []
    for i in range(logn):
        from schemas.polynomial_jax import apply_automorphism
        sa = torch.as_tensor(apply_automorphism(S_np, 2 ** (logn - i) + 1), device=dev)
        A, b = _encrypt_rlwe_dev(_glev_dev(-sa, lbp, lp), S, dpp["sigma"], gen)
        gks.append(torch.stack([A, b], dim=1))                                      # (lp, 2, N)
    gk = torus.poly_fft(torch.stack(gks))
    # packing key (RLWE_jax.py:205-214)
    lbs, ls = int(np.log2(dp["beta_ks"]) + 1e-6), dp["l_ks"]
    poly = torch.zeros(N, N, dtype=torch.int64, device=dev)
    poly[:, 0] = S
    msg = _glev_dev(poly, lbs, ls).permute(1, 0, 2).contiguous()                   # (N, ls, N)
    A, b = _encrypt_rlwe_dev(msg, S, dp["sigma"], gen)
    ksp = torus.poly_fft(torch.stack([A, b], dim=2))
    return {"bsk": bsk, "ks": ks, "gk": gk, "ksp": ksp}, (S_np, s.cpu().numpy())
