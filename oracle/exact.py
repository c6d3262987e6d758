"""ctypes wrapper of the exact-integer C oracle (oracle/c/tfhe_exact.c).

TEST INFRASTRUCTURE ONLY (see oracle/tfhe_oracle.py).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtfhe_exact.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        lib.tfhe_exact_blind_rotate.argtypes = [P, ctypes.c_long, P, P, P, I, I, I, I, I, I, I]
        lib.tfhe_exact_blind_rotate.restype = None
        lib.tfhe_exact_negacyclic_mul.argtypes = [P, P, P, I, I]
        lib.tfhe_exact_negacyclic_mul.restype = None
        _lib = lib
    return _lib


def negacyclic_mul(a, b, log_q: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=np.int64)
    out = np.empty_like(a)
    _load().tfhe_exact_negacyclic_mul(a.ctypes.data, b.ctypes.data, out.ctypes.data,
                                      a.shape[-1], log_q)
    return out


def blind_rotate(ct, bsk_coef_k, idx, log_beta: int, log_q: int, L: int) -> np.ndarray:
    """Exact collapsed blind rotation.

    ct (B,2,N) or (2,N) int64 (broadcast), bsk_coef_k (n_iter,2,L,2,m,N) int64,
    idx (B,n_iter,m) int32 -> (B,2,N) int64 centred mod q.
    """
    ct = np.ascontiguousarray(ct, dtype=np.int64)
    bsk = np.ascontiguousarray(bsk_coef_k, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    B, n_iter, m = idx.shape
    N = ct.shape[-1]
    stride = 0 if ct.ndim == 2 else 2 * N
    out = np.empty((B, 2, N), np.int64)
    _load().tfhe_exact_blind_rotate(ct.ctypes.data, stride, bsk.ctypes.data, idx.ctypes.data,
                                    out.ctypes.data, B, n_iter, m, N, L, log_beta, log_q)
    return out
