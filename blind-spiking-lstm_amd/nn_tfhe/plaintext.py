"""Plaintext semantics of the encrypted spiking-LSTM layer and MLP head.

What CipherSpikeLSTM and CipherMLP compute, restated on plaintext integers so
that callers can check decrypted results (the role of the reference's
compare_cipher_plain.py:17-19, which compares cipher and plain predictions).
It follows the plaintext model's recurrence (nn_jax/Spike_LSTM.py:9-157):
  C_t = C_(t-1) * 1[f >= gamma_f] + eta * 1[i >= gamma_i]
  H_t = C_t * 1[o >= gamma_o]
with every gate read through the step LUT of nn_tfhe/utils_spike.py:72-87
(pre-activations in [-s, s), one threshold per many-LUT slot k and group g,
unit u = k*G + g) and, beyond [-s, s), the LUT's negacyclic extension: the
blind rotation reads the test polynomial mod X^N + 1, so a pre-activation
p + 2s reads -T(p) (period 4s).  Values live on the torus in units of beta_x
(message = value * beta_x mod q), so they are kept mod q / beta_x.

A unit is *determined* when every pre-activation it reads lies at least
`margin` from each discontinuity of the extended LUT (its threshold, the
range ends +-s and their images); elsewhere the encrypted result may take
either neighbouring value (noise, LUT grid), and checks skip it.
"""
from __future__ import annotations

import numpy as np


def fold(pre, s: float):
    """Pre-activation reduced to [-2s, 2s) (the torus period 4s)."""
    return np.mod(np.asarray(pre, np.float64) + 2 * s, 4 * s) - 2 * s


def step_lut(pre, gamma, s: float):
    """1[p >= gamma] on [-s, s) and its negacyclic extension, p = fold(pre):
    T(p) for p in [-s, s), -T(p - 2s) above, -T(p + 2s) below."""
    p = fold(pre, s)
    inside = (p >= -s) & (p < s)
    shifted = np.where(p >= s, p - 2 * s, p + 2 * s)
    return np.where(inside, (p >= gamma).astype(np.int64), -(shifted >= gamma).astype(np.int64))


def lut_margin(pre, gamma, s: float):
    """Distance of fold(pre) to the nearest discontinuity of step_lut."""
    p = fold(pre, s)
    pts = [gamma, gamma - 2 * s, gamma + 2 * s, np.full_like(np.asarray(gamma, np.float64), s),
           np.full_like(np.asarray(gamma, np.float64), -s),
           np.full_like(np.asarray(gamma, np.float64), 2 * s),
           np.full_like(np.asarray(gamma, np.float64), -2 * s)]
    return np.min(np.stack([np.abs(p - np.asarray(q, np.float64)) for q in pts]), axis=0)


def wrap(v, modulus: int):
    """Centred representative mod `modulus` (the torus in units of beta_x)."""
    v = np.asarray(v, np.int64)
    return np.mod(v + modulus // 2, modulus) - modulus // 2


class PlainSpikingLSTM:
    """One layer: weights W[c] (input_dim + hidden, G) integers and biases
    b[c] (G,) for the gates c in 'fio', in units where the pre-activation is
    (hx @ W + b) / w_scale; thresholds gamma[c] (n_lut, G); eta (n_lut, G);
    LUT half-range s.  modulus = q / beta_x (the values' torus)."""

    def __init__(self, W: dict, b: dict, gamma: dict, eta, s: float, n_lut: int, w_scale: float,
                 modulus: int):
        self.W = {c: np.asarray(W[c], np.int64) for c in "fio"}
        self.b = {c: np.asarray(b[c], np.float64) for c in "fio"}
        self.gamma = {c: np.asarray(gamma[c], np.float64) for c in "fio"}
        self.eta = np.asarray(eta, np.int64)
        self.s, self.n_lut, self.w_scale, self.modulus = s, n_lut, w_scale, modulus
        self.G = self.W["f"].shape[1]
        self.hidden = self.G * n_lut

    def pre(self, x, H):
        """Gate pre-activations {c: (..., G)} for spikes x (..., input_dim)
        and hidden state H (..., hidden) (integers mod `modulus`)."""
        hx = np.concatenate([np.asarray(x, np.int64), np.asarray(H, np.int64)], axis=-1)
        return {c: (hx @ self.W[c] + self.b[c]) / self.w_scale for c in "fio"}

    def step(self, x, H, C, margin: float):
        """(H_t, C_t, determined) for one timestep; determined (..., hidden)
        marks the units whose three gates are at least `margin` from every
        LUT discontinuity."""
        pre = self.pre(x, H)
        lead = np.asarray(x).shape[:-1]
        Hn = np.zeros(lead + (self.hidden,), np.int64)
        Cn = np.zeros_like(Hn)
        det = np.zeros(lead + (self.hidden,), bool)
        for k in range(self.n_lut):
            sl = slice(k * self.G, (k + 1) * self.G)
            gate = {c: step_lut(pre[c], self.gamma[c][k], self.s) for c in "fio"}
            ok = np.ones(lead + (self.G,), bool)
            for c in "fio":
                ok &= lut_margin(pre[c], self.gamma[c][k], self.s) >= margin
            c_new = wrap(np.asarray(C, np.int64)[..., sl] * gate["f"] + self.eta[k] * gate["i"],
                         self.modulus)
            Cn[..., sl] = c_new
            Hn[..., sl] = wrap(c_new * gate["o"], self.modulus)
            det[..., sl] = ok
        return Hn, Cn, det


def relu_ramp(pre, s: float, n_entries: int):
    """The ReLU head's ramp LUT (nn_tfhe/activation.py:195-201): pre on
    [0, s), 0 on [-s, 0), quantised to the LUT grid of n_entries = N/2 steps
    over [0, s); beyond [-s, s) its negacyclic extension (0 on [s, 2s),
    -ramp(p + 2s) on [-2s, -s)).  Discontinuous at +-s only."""
    p = fold(pre, s)
    ramp = lambda v: np.floor(v * n_entries / s + 0.5) * s / n_entries
    return np.where((p >= 0) & (p < s), ramp(p),
                    np.where(p < -s, -ramp(p + 2 * s), 0.0))


def lut_index_noise(dict_params_lut: dict, dict_params_ks_LWE: dict, degree: int, q: float,
                    n_lut: int) -> float:
    """Standard deviation, in LUT entries (q * n_lut / 2N of phase each), of
    the LUT index a programmable bootstrap reads at these parameters: the
    modulus switch rounds each of the n/c group sums and the body to the
    grid (uniform, +-1/2 entry each: variance (n/c + 1)/12), and the key
    switch adds sqrt(N * l_ks) * (beta_ks / 2) * sigma_lwe (the
    key-switching key's noise times digits of at most beta_ks/2).  Units
    within a few of these of a LUT discontinuity may read either side."""
    n_groups = dict_params_ks_LWE["degree"] // dict_params_lut["collapse"] + 1
    entry = q * n_lut / (2 * dict_params_lut["degree"])
    ks = (np.sqrt(degree * dict_params_ks_LWE["l_ks"]) * (dict_params_ks_LWE["beta_ks"] / 2)
          * dict_params_ks_LWE["sigma"] / entry)
    return float(np.sqrt(n_groups / 12 + ks ** 2))
