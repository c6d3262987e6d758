"""ctypes binding of libbslstm_tfhe.so (C ABI declared in include/bslstm_tfhe.h).

Plays the role of the reference's nanobind module `cuTFHE._core`
(csrc/bindings.cu:20-31): it is the only place the native library is loaded.
There is no CPU fallback: every entry point needs the HIP library and a GPU,
and loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import torch

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libbslstm_tfhe.so")

BSLSTM_OK = 0
BSLSTM_INVALID_ARGUMENT = 1
BSLSTM_INTERNAL = 2

_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64

SIGNATURES = {
    "bslstm_last_error": ([], ctypes.c_char_p),
    "bslstm_abi_version": ([], ctypes.c_int),
    "bslstm_mul_bootstrap": ([_P, _P, _I64, _P, _P, _P, _I64, _I64, _P, _I64, _P, _I64,
                              _I64, _I64, _I64, _I64], ctypes.c_int),
    "bslstm_mul_bootstrap_batched": ([_P, _P, _I64, _P, _P, _P, _I64, _I64, _P, _I64, _P, _I64,
                                      _I64, _I64, _I64, _I64], ctypes.c_int),
    "bslstm_mul_bootstrap_rows": ([_P, _P, _I64, _P, _P, _P, _I64, _I64, _P, _I64, _P, _I64,
                                   _I64, _I64, _I64, _I64], ctypes.c_int),
    "bslstm_mul_bootstrap_pair_workspace_bytes": ([_I64, _I64, _I64], ctypes.c_int64),
    "bslstm_mul_bootstrap_pair": ([_P, _P, _I64, _P, _P, _P, _I64, _I64, _P, _I64, _P, _I64,
                                   _I64, _I64, _I64, _I64, _P, _I64, _I64], ctypes.c_int),
    "bslstm_mul_bootstrap_pair_slots": ([_P, _P, _I64, _P, _P, _P, _I64, _P, _I64, _P, _I64,
                                         _I64, _I64, _I64, _I64, _P, _I64, _I64], ctypes.c_int),
    "bslstm_rotation_indices": ([_P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _P], ctypes.c_int),
    "bslstm_rotate_extract": ([_P, _P, _P, _I64, _I64, _I64, _I64, _P, _P], ctypes.c_int),
    "bslstm_key_switch_lwe": ([_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _P, _P],
                              ctypes.c_int),
    "bslstm_rotation_indices_i64": ([_P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _P],
                                    ctypes.c_int),
    "bslstm_rotate_extract_i64": ([_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P, _P],
                                  ctypes.c_int),
    "bslstm_key_switch_lwe_i64": ([_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _P,
                                   _P], ctypes.c_int),
    "bslstm_linear_extract": ([_P, _P, _I64, _P, _P, _I64, _P, _I64, _I64, _I64, _I64, _P, _P],
                              ctypes.c_int),
    "bslstm_pack_lwe_to_rlwe": ([_P, _P, _P, _I64, _P, _I64, _I64, _I64, _I64, _P],
                                ctypes.c_int),
    "bslstm_lut_product_sum": ([_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P], ctypes.c_int),
    "bslstm_poly_fft": ([_P, _P, _I64, _I64, _P], ctypes.c_int),
    "bslstm_poly_ifft": ([_P, _P, _I64, _I64, _P], ctypes.c_int),
    "bslstm_packing_workspace_bytes": ([_I64, _I64, _I64], ctypes.c_int64),
    "bslstm_packing": ([_P, _P, _P, _I64, _I64, _I64, _P, _I64, _I64, _I64, _I64, _P, _I64, _P],
                       ctypes.c_int),
    "bslstm_torus_axpy": ([_P, _P, _P, _I64, _I64, _I64, _I64, _P], ctypes.c_int),
    "bslstm_rotate_poly": ([_P, _P, _I64, _I64, _P, _I64, _I64, _P], ctypes.c_int),
    "bslstm_ks_mfma_key_bytes": ([_I64, _I64, _I64], ctypes.c_int64),
    "bslstm_ks_mfma_prepare_key": ([_P, _P, _P, _I64, _I64, _I64, _P], ctypes.c_int),
    "bslstm_ks_mfma_workspace_bytes": ([_I64, _I64, _I64], ctypes.c_int64),
    "bslstm_key_switch_lwe_mfma": ([_P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _P, _I64,
                                    _P, _P], ctypes.c_int),
    "bslstm_key_switch_lwe_mfma_noalloc_max_bytes": ([], ctypes.c_int64),
    "bslstm_key_switch_lwe_mfma_noalloc": ([_P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64,
                                            _P, _P], ctypes.c_int),
    "bslstm_mul_bootstrap_split_key_bytes": ([_P, _I64], ctypes.c_int64),
    "bslstm_mul_bootstrap_split_prepare_key": ([_P, _P, _P, _I64, _I64, _P], ctypes.c_int),
    "bslstm_mul_bootstrap_split_workspace_bytes": ([_I64, _P, _I64, _I64], ctypes.c_int64),
    "bslstm_mul_bootstrap_split": ([_P, _P, _I64, _P, _P, _P, _I64, _P, _I64, _I64, _I64, _I64,
                                    _I64, _P, _I64], ctypes.c_int),
    "bslstm_mul_bootstrap_slots": ([_P, _P, _I64, _P, _P, _P, _I64, _P, _I64, _P, _I64, _I64, _I64,
                                    _I64, _I64, _I64], ctypes.c_int),
    "bslstm_key_switch_lwe_mfma_slots": ([_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64,
                                          _I64, _P, _I64, _P, _P], ctypes.c_int),
    "bslstm_pack_lwe_to_rlwe_slots": ([_P, _P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _I64, _P],
                                      ctypes.c_int),
    "bslstm_packing_slots": ([_P, _P, _P, _I64, _I64, _I64, _P, _P, _I64, _I64, _I64, _I64, _I64,
                              _P, _I64, _P], ctypes.c_int),
}


class TFHEInvalidArgument(ValueError):
    """ffi::Error::InvalidArgument of the reference handler."""


class TFHEInternalError(RuntimeError):
    """ffi::Error::Internal of the reference handler (launch failure)."""


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python blind-spiking-lstm_amd/build_lib.py`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status == BSLSTM_OK:
        return
    msg = lib().bslstm_last_error().decode()
    if status == BSLSTM_INVALID_ARGUMENT:
        raise TFHEInvalidArgument(msg)
    raise TFHEInternalError(msg)


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def require_device(*tensors: torch.Tensor) -> torch.device:
    dev = None
    for t in tensors:
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TFHEInvalidArgument("operands must be torch tensors on a ROCm device")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise TFHEInvalidArgument("operands live on different devices")
    return dev
