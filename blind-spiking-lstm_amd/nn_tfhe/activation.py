"""Bootstrapper classes and key preparation of the reference's
nn_tfhe/activation.py, on the GPU engine.

Keys are generated on the host and moved to the GPU once, in the layouts the
kernels read: the bootstrapping key already reshaped to
(n_iter, 2, l, 2, 2^c, N/2) complex128 and the key-switching key as float64
(l, N_in, n) / (l, N_in, 1).
"""
from __future__ import annotations

import copy
from typing import Callable

import numpy as np
import torch

from cuTFHE.multiply import log2_int, prepare_bsk
from schemas.LWE_jax import get_keyswitch_sk_LWE, key_switch_LWE, rng_of
from schemas.RLWE_jax import (bootstrapping, bsk_to_fourier, cuboot_merge,
                              get_boostrapping_key_collapse)


def prepare_bootstrapping(key, sk, sk_lut, sk_bsk, dict_params: dict, dict_params_lut: dict,
                          dict_params_ks_LWE: dict, function_lut: Callable, factor, collapse=None):
    """nn_tfhe/activation.py:15-67.

    Returns (key_switching_key_back, ksk_packing, bsk_ordered,
    key_switching_key_bs, encrypted_LUT) like the reference.  The packing
    key-switching key belongs to LWE->RLWE packing, which this engine does not
    implement yet; ksk_packing is returned as None.
    """
    if collapse is None:
        raise NotImplementedError("only the collapsed bootstrapping key is live in the reference "
                                  "(nn_tfhe/activation.py:44-47)")
    rng = rng_of(key)
    t = dict_params["t"]
    degree_lut = dict_params_lut["degree"]
    key_switching_key_bs = get_keyswitch_sk_LWE(sk_bsk, sk, dict_params_ks_LWE, rng)
    LUT = function_lut(t, degree_lut)
    encrypted_LUT = [np.zeros(degree_lut), np.asarray(LUT, np.float64) * factor]
    BSK = get_boostrapping_key_collapse(sk_lut, sk_bsk, dict_params_lut, rng, collapse)
    bsk_ordered = bsk_to_fourier(BSK)
    dict_params_copy = copy.deepcopy(dict_params)
    dict_params_copy["beta_ks"] = dict_params_lut["beta_ks"]
    dict_params_copy["l_ks"] = dict_params_lut["l_ks"]
    key_switching_key_back = get_keyswitch_sk_LWE(sk, sk_lut, dict_params_copy, rng)
    return key_switching_key_back, None, bsk_ordered, key_switching_key_bs, encrypted_LUT


def _to_device(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device)
    return torch.as_tensor(np.asarray(x), device=device)


class HeavysideBoostrapper:
    """nn_tfhe/activation.py:70-162: key switch to the small LWE key, then
    many-LUT programmable bootstrap."""

    def __init__(self, bootstrapping_key, key_switching_key_bs, ksk_packing, galois_key,
                 dict_params, dict_params_lut, dict_params_ks_LWE, dict_params_packing,
                 collapse, all_rot_possible_fourier, n_lut, device="cuda"):
        self.device = torch.device(device)
        n = dict_params_ks_LWE["degree"]
        self.bootstrapping_key = prepare_bsk(_to_device(bootstrapping_key, self.device), n, collapse)
        self.key_switching_key_bs = tuple(_to_device(k, self.device) for k in key_switching_key_bs)
        self.ksk_packing = ksk_packing
        self.galois_key = galois_key
        self.dict_params = dict_params
        self.dict_params_lut = dict_params_lut
        self.dict_params_ks_LWE = dict_params_ks_LWE
        self.dict_params_packing = dict_params_packing
        self.collapse = collapse
        # the FFT of X^r is evaluated in-kernel; the table is only shape-checked
        self.all_rot_possible_fourier = (None if all_rot_possible_fourier is None else
                                         _to_device(all_rot_possible_fourier, self.device))
        self.log2q = log2_int(dict_params_lut["q"])
        self.log2beta = log2_int(dict_params_lut["beta_bs"])
        self.n_lut = n_lut

    def __call__(self, h, encrypted_LUT):
        c_lwe_ks = key_switch_LWE(self.key_switching_key_bs, h, self.dict_params_ks_LWE)
        p = self.dict_params_lut
        return bootstrapping(c_lwe_ks, encrypted_LUT, self.bootstrapping_key, p["q"], p["beta_bs"],
                             p["l_bs"], p["degree"], self.collapse, self.all_rot_possible_fourier)

    def cuboot(self, h, encrypted_LUT):
        c_lwe_ks = key_switch_LWE(self.key_switching_key_bs, h, self.dict_params_ks_LWE)
        p = self.dict_params_lut
        return cuboot_merge(c_lwe_ks, encrypted_LUT, self.bootstrapping_key, p["q"], p["beta_bs"],
                            p["l_bs"], p["degree"], self.collapse, self.all_rot_possible_fourier,
                            self.n_lut)
