"""Cryptographic parameters of the encrypted spiking LSTM.

The constants and dictionaries of the reference's parameters_crypto.py:6-84
(the interface its callers read: dict_params, dict_params_lut,
dict_params_ks_LWE, dict_params_packing, beta_x, beta_w).  `make_params`
builds the same four dictionaries for other sizes (tests use smaller rings).
The reference also builds all_rot_possible_fourier (parameters_crypto.py:86-91,
the transform of X^r for r in [0, 2N)); the kernels evaluate those roots
in-kernel, so it is None here.
"""
from __future__ import annotations


def make_params(degree: int = 2 ** 11, n_ks: int = 702, log_q: int = 64,
                sigma: float = 2 ** 17.0, sigma_lut: float = 2 ** 17.0,
                sigma_lwe: float = 2 ** 48.6,
                beta_ks_monomial: int = 2 ** 2, l_ks_monomial: int = 20,
                beta_ks_default: int = 2 ** 8, l_ks_default: int = 4,
                beta_ks_lwe: int = 2, l_ks_lwe: int = 20,
                beta_bs: int = 2 ** 13, l_bs: int = 2, collapse: int = 3):
    """The four parameter dictionaries of parameters_crypto.py:36-84."""
    q = float(2 ** log_q)
    t = degree
    dict_params = {"t": t, "degree": degree, "q": q, "beta_bs": beta_bs, "l_bs": l_bs,
                   "beta_ks": beta_ks_default, "l_ks": l_ks_default, "sigma": sigma}
    dict_params_ks_LWE = {"t": t, "degree": n_ks, "q": q, "beta_bs": beta_bs, "l_bs": l_bs,
                          "beta_ks": beta_ks_lwe, "l_ks": l_ks_lwe, "sigma": sigma_lwe}
    dict_params_packing = {"t": t, "degree": degree, "q": q, "beta_bs": beta_bs, "l_bs": l_bs,
                           "beta_ks": beta_ks_monomial, "l_ks": l_ks_monomial, "sigma": sigma}
    dict_params_lut = {"t": degree, "degree": degree, "q": q, "beta_bs": beta_bs, "l_bs": l_bs,
                       "beta_ks": beta_ks_default, "l_ks": l_ks_default, "collapse": collapse,
                       "sigma": sigma_lut}
    return dict_params, dict_params_lut, dict_params_ks_LWE, dict_params_packing


# parameters_crypto.py:6-33
q = float(2 ** 64)
Q = q
degree_lut = 2 ** 11
t = degree_lut
degree = 2 ** 11
n_keyswitch_bootstrapping = 702
sigma = 2 ** 17.0
sigma_lut = 2 ** 17.0
sigma_lwe = 2 ** 48.6
beta_ks_monomial, l_ks_monomial = 2 ** 2, 20
beta_ks_default, l_ks_default = 2 ** 8, 4
beta_ks_lwe, l_ks_lwe = 2 * 1, 20
beta_bs, l_bs = 2 ** 13, 2
collapse = 3

dict_params, dict_params_lut, dict_params_ks_LWE, dict_params_packing = make_params()

# parameters_crypto.py:79-80
beta_x = float(2 ** 56)
beta_w = q / (beta_x * 4)

all_rot_possible_fourier = None
