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
