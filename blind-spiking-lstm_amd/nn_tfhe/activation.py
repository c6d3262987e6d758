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

from cuTFHE import torus
from cuTFHE.multiply import log2_int, prepare_bsk
from schemas.LWE_jax import get_keyswitch_sk_LWE, key_switch_LWE, rng_of
from schemas.RLWE_jax import (bootstrapping, bsk_to_fourier, cuboot_merge,
                              get_boostrapping_key_collapse, get_packing_KSK)
from schemas.text_jax import multiply_plaintext_ciphertext, pack_lwe_to_rlwe, packing, prepare_ksp_fft


def prepare_bootstrapping(key, sk, sk_lut, sk_bsk, dict_params: dict, dict_params_lut: dict,
                          dict_params_ks_LWE: dict, function_lut: Callable, factor, collapse=None):
    """nn_tfhe/activation.py:15-67.

    Returns (key_switching_key_back, ksk_packing, bsk_ordered,
    key_switching_key_bs, encrypted_LUT) like the reference.
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
    ksk_packing = get_packing_KSK(sk, sk, dict_params_copy, dict_params_copy["q"],
                                  dict_params_copy["beta_ks"], dict_params_copy["l_ks"], rng)
    key_switching_key_back = get_keyswitch_sk_LWE(sk, sk_lut, dict_params_copy, rng)
    return key_switching_key_back, ksk_packing, bsk_ordered, key_switching_key_bs, encrypted_LUT


def _to_device(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device)
    return torch.as_tensor(np.asarray(x), device=device)


def _ks_key(key, dict_params_ks: dict, device):
    """The LWE key-switching key on the device; an int64 key with digits of
    at most 7 bits also gets its int8 limb tiles for the matrix-core key
    switch (a third tuple entry that key_switch_LWE picks up)."""
    ka, kb = (_to_device(k, device) for k in key[:2])
    if ka.dtype == torch.int64 and log2_int(dict_params_ks["beta_ks"]) <= 7:
        kb2 = kb.reshape(ka.shape[0], ka.shape[1])
        return ka, kb, torus.prepare_ks_key(ka, kb2)
    return ka, kb


class HeavysideBoostrapper:
    """nn_tfhe/activation.py:70-162: key switch to the small LWE key, then
    many-LUT programmable bootstrap."""

    def __init__(self, bootstrapping_key, key_switching_key_bs, ksk_packing, galois_key,
                 dict_params, dict_params_lut, dict_params_ks_LWE, dict_params_packing,
                 collapse, all_rot_possible_fourier, n_lut, device="cuda"):
        self.device = torch.device(device)
        n = dict_params_ks_LWE["degree"]
        self.bootstrapping_key = prepare_bsk(_to_device(bootstrapping_key, self.device), n, collapse)
        self.key_switching_key_bs = _ks_key(key_switching_key_bs, dict_params_ks_LWE, self.device)
        # packing key in the transform layout the packing kernel reads, once
        self.ksk_packing = (None if ksk_packing is None else
                            prepare_ksp_fft(ksk_packing if isinstance(ksk_packing, torch.Tensor)
                                            else tuple(_to_device(k, self.device)
                                                       for k in ksk_packing)))
        self.galois_key = (None if galois_key is None else
                           _to_device(galois_key, self.device).to(torch.complex128))
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
        """activation.py:120-146: key switch N -> n, then one bootstrap per LWE
        with the encrypted test polynomial (one LUT)."""
        c_lwe_ks = key_switch_LWE(self.key_switching_key_bs, h, self.dict_params_ks_LWE)
        p = self.dict_params_lut
        return bootstrapping(c_lwe_ks, encrypted_LUT, self.bootstrapping_key, p["q"], p["beta_bs"],
                             p["l_bs"], p["degree"], self.collapse, self.all_rot_possible_fourier)

    def cuboot(self, h, encrypted_LUT):
        return self.cuboot_switched(self.key_switch(h), encrypted_LUT)

    def key_switch(self, h):
        """The key switch cuboot starts with (activation.py:150)."""
        return key_switch_LWE(self.key_switching_key_bs, h, self.dict_params_ks_LWE)

    def cuboot_switched(self, c_lwe_ks, encrypted_LUT):
        """cuboot after its key switch (activation.py:151-162): blind
        rotation of the LUT and many-LUT extraction."""
        p = self.dict_params_lut
        return cuboot_merge(c_lwe_ks, encrypted_LUT, self.bootstrapping_key, p["q"], p["beta_bs"],
                            p["l_bs"], p["degree"], self.collapse, self.all_rot_possible_fourier,
                            self.n_lut)

    def prepare_LUT(self, c_lwe, input_lut):
        """activation.py:103-106: trace-pack the LWE c_lwe = (a (..., N), b (..., 1))
        into an RLWE and multiply it by the plaintext LUT polynomial
        input_lut (..., N).  Returns (A (..., N), B (..., N))."""
        p = self.dict_params_packing
        c_rlwe = pack_lwe_to_rlwe(c_lwe, self.galois_key, p["beta_ks"], p["l_ks"], p["q"],
                                  p["degree"])
        return multiply_plaintext_ciphertext(input_lut, c_rlwe, p["degree"], p["q"])

    def lut_fft(self, lut) -> torch.Tensor:
        """Transform of plaintext LUT polynomials (..., N) for prepare_LUT_sum."""
        return torus.poly_fft(torus._i64(_to_device(lut, self.device)))

    def prepare_LUT_sum(self, c_lwe, lut_fft):
        """The LUT preparation of Spike_LSTM.py:129-132 in one pass:
        vmap(vmap(prepare_LUT)) over (n_lut, G) followed by the sum over the
        n_lut slots.  c_lwe = (masks (S, n_lut, G, N), bodies (S, n_lut, G[, 1]));
        lut_fft = self.lut_fft(lut) of the (n_lut, G, N) LUT polynomials.
        Returns RLWEs int64 (S, G, 2, N)."""
        p = self.dict_params_packing
        masks, bodies = c_lwe
        rl = torus.pack_lwe_to_rlwe(masks, bodies.reshape(masks.shape[:-1]), self.galois_key,
                                    log2_int(p["beta_ks"]), p["l_ks"], log2_int(p["q"]))
        return torus.lut_product_sum(rl, lut_fft, masks.shape[-3], log2_int(p["q"]))

    def packing(self, x):
        """activation.py:109-115: pack LWEs (masks (..., n_lwes, N), bodies
        (..., n_lwes, 1)) into RLWEs (A, B) with the packing key."""
        p = self.dict_params
        return packing(x, self.ksk_packing, p["q"], p["beta_ks"], p["l_ks"], p["degree"])


class ReluBoostrapper:
    """nn_tfhe/activation.py:166-238: key switch, then bootstrap every hidden
    unit through its own ramp LUT (the MLP head of CipherMLP)."""

    def __init__(self, bootstrapping_key, key_switching_key_bs, ksk_packing, dict_params,
                 dict_params_lut, dict_params_ks_LWE, dict_params_packing, collapse,
                 all_rot_possible_fourier, beta_x, s_head, device="cuda"):
        self.device = torch.device(device)
        n = dict_params_ks_LWE["degree"]
        self.bootstrapping_key = prepare_bsk(_to_device(bootstrapping_key, self.device), n, collapse)
        self.key_switching_key_bs = _ks_key(key_switching_key_bs, dict_params_ks_LWE, self.device)
        self.ksk_packing = (None if ksk_packing is None else
                            prepare_ksp_fft(ksk_packing if isinstance(ksk_packing, torch.Tensor)
                                            else tuple(_to_device(k, self.device)
                                                       for k in ksk_packing)))
        self.dict_params = dict_params
        self.dict_params_lut = dict_params_lut
        self.dict_params_ks_LWE = dict_params_ks_LWE
        self.dict_params_packing = dict_params_packing
        self.collapse = collapse
        self.all_rot_possible_fourier = (None if all_rot_possible_fourier is None else
                                         _to_device(all_rot_possible_fourier, self.device))
        self.beta_x = beta_x
        s_head = np.asarray(s_head, np.float64).reshape(-1)
        self.hidden_dim = s_head.shape[0]
        N = dict_params_lut["degree"]
        # activation.py:195-201: ramp over the first half, zero over the second
        ramps = np.stack([np.concatenate([np.arange(N // 2) * (s * 2 / N) * beta_x,
                                          np.zeros(N // 2)]) for s in s_head])
        body = np.rint(ramps).astype(np.int64)
        self.LUT = [torch.zeros((self.hidden_dim, N), dtype=torch.int64, device=self.device),
                    torch.as_tensor(body, device=self.device)]

    def __call__(self, h):
        """h = (masks (..., hidden, N), bodies (..., hidden, 1)): unit u goes
        through LUT u; leading axes (sequences) share the LUTs."""
        masks, bodies = h
        lead, N = masks.shape[:-2], masks.shape[-1]
        hid = masks.shape[-2]
        reps = masks.numel() // (hid * N)
        c_lwe_ks = key_switch_LWE(self.key_switching_key_bs,
                                  (masks.reshape(-1, N), bodies.reshape(-1, 1)),
                                  self.dict_params_ks_LWE)
        lut = [x.unsqueeze(0).expand(reps, hid, x.shape[-1]).reshape(reps * hid, -1)
               for x in self.LUT]
        p = self.dict_params_lut
        om, ob = cuboot_merge(c_lwe_ks, lut, self.bootstrapping_key, p["q"], p["beta_bs"],
                              p["l_bs"], p["degree"], self.collapse, self.all_rot_possible_fourier,
                              None)
        return om.reshape(*lead, hid, -1), ob.reshape(*lead, hid, 1)

    def packing(self, x):
        p = self.dict_params
        return packing(x, self.ksk_packing, p["q"], p["beta_ks"], p["l_ks"], p["degree"])
