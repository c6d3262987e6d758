"""The C-ABI library loads, exports every symbol include/bslstm_tfhe.h
declares, and rejects bad arguments on the host before any GPU work (these
error paths run without a GPU: validation precedes every HIP call)."""
import ctypes
import os
import re

import pytest

from cuTFHE import _core

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "bslstm_tfhe.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(bslstm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert {"bslstm_mul_bootstrap", "bslstm_rotation_indices", "bslstm_rotate_extract",
            "bslstm_key_switch_lwe", "bslstm_last_error", "bslstm_abi_version"} <= set(syms)


def test_library_exports_every_declared_symbol():
    lib = _core.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert name in _core.SIGNATURES, f"{name} has no ctypes signature"


def test_abi_version():
    assert _core.lib().bslstm_abi_version() == 2


def _mul(degree=2048, l=3, dims=(250, 2, 3, 2, 8, 1024), ct_elems=None, idx_elems=None, batch=4,
         log2_beta=7, log2_q=32):
    arr = (ctypes.c_int64 * 6)(*dims)
    ct_elems = 2 * degree if ct_elems is None else ct_elems
    idx_elems = batch * dims[0] * dims[4] if idx_elems is None else idx_elems
    return _core.lib().bslstm_mul_bootstrap(None, 8, ct_elems, 8, arr, None, 0, 0, 8, idx_elems, 8,
                                            batch, log2_beta, log2_q, l, degree)


@pytest.mark.parametrize("kw,msg", [
    (dict(degree=100), "Unsupported degree"),
    (dict(degree=8192), "Unsupported degree"),
    (dict(l=5, dims=(250, 2, 5, 2, 8, 1024)), "Unsupported l"),
    (dict(l=0), "Unsupported l"),
    (dict(dims=(250, 2, 2, 2, 8, 1024)), "bsk trailing dims"),
    (dict(dims=(250, 2, 3, 2, 8, 512)), "bsk trailing dims"),
    (dict(ct_elems=3 * 2048), "ct_init has"),
    (dict(idx_elems=7), "indices has"),
    (dict(log2_beta=11), "log2_beta"),
])
def test_mul_bootstrap_rejects(kw, msg):
    assert _mul(**kw) == _core.BSLSTM_INVALID_ARGUMENT
    assert msg in _core.lib().bslstm_last_error().decode()


def test_other_entry_points_reject():
    lib = _core.lib()
    assert lib.bslstm_rotation_indices(None, 8, 4, 10, 3, 32, 2048, 4, 8) == _core.BSLSTM_INVALID_ARGUMENT
    assert lib.bslstm_rotate_extract(None, 8, 8, 4, 1000, 32, 4, 8, 8) == _core.BSLSTM_INVALID_ARGUMENT
    assert lib.bslstm_key_switch_lwe(None, 8, 8, 8, 8, 4, 2048, 750, 4, 5, 64, 8, 8) == \
        _core.BSLSTM_INVALID_ARGUMENT
    assert "log2_q <= 32" in lib.bslstm_last_error().decode()


def test_empty_batch_is_noop():
    assert _mul(batch=0, idx_elems=0) == _core.BSLSTM_OK


def test_python_surface_raises_invalid_argument():
    import torch
    from cuTFHE import multiply
    with pytest.raises(_core.TFHEInvalidArgument):
        multiply.rotation_indices(torch.zeros(3, 10, dtype=torch.float64), 32, 2048, 3, 1)


def test_every_bound_symbol_is_declared():
    assert set(_core.SIGNATURES) == set(declared_symbols())


def test_lstm_path_entry_points_reject_on_host():
    """Validation of the LSTM-path entry points runs before any GPU work."""
    lib = _core.lib()
    bad = _core.BSLSTM_INVALID_ARGUMENT
    # Linear: unsupported degree, K > degree
    assert lib.bslstm_linear_extract(None, 8, 2 * 100, 8, 8, 4, None, 1, 4, 100, 64, 8, 8) == bad
    assert lib.bslstm_linear_extract(None, 8, 2 * 256, 8, 8, 300, None, 1, 4, 256, 64, 8, 8) == bad
    # trace packing: l * log2_beta > log2_q
    assert lib.bslstm_pack_lwe_to_rlwe(None, 8, 8, 3, 8, 2048, 4, 20, 64, 8) == bad
    assert "l*log2_beta" in lib.bslstm_last_error().decode()
    # packing: workspace too small, n_lwes > degree
    assert lib.bslstm_packing(None, 8, 8, 1, 40, 256, 8, 256, 8, 4, 64, 8, 16, 8) == bad
    assert "workspace" in lib.bslstm_last_error().decode()
    assert lib.bslstm_packing(None, 8, 8, 1, 300, 256, 8, 256, 8, 4, 64, 8, 1 << 30, 8) == bad
    # MFMA key switch: digits wider than int8
    assert lib.bslstm_key_switch_lwe_mfma(None, 8, 8, 8, 4, 2048, 702, 8, 4, 64, 8, 1 << 30, 8,
                                          16) == bad
    assert "log2_beta_ks <= 7" in lib.bslstm_last_error().decode()
    # key-slot forms need slots
    dims = (_core._I64 * 6)(234, 2, 2, 2, 8, 1024)
    assert lib.bslstm_mul_bootstrap_slots(None, 8, 4096, 8, dims, None, 4, 8, 4 * 234 * 8, 8, 4,
                                          13, 64, 2, 2048, 2) == bad
    assert lib.bslstm_packing_slots(None, 8, 8, 1, 40, 256, 8, None, 2, 256, 8, 4, 64, 8, 1 << 30,
                                    8) == bad
    # unknown blind-rotation form
    assert lib.bslstm_mul_bootstrap_slots(None, 8, 4096, 8, dims, 8, 4, 8, 4 * 234 * 8, 8, 4,
                                          13, 64, 2, 2048, 4) == bad
    assert "form must be" in lib.bslstm_last_error().decode()
    # aliasing the key-switch output onto its input is refused
    assert lib.bslstm_key_switch_lwe_i64(None, 8, 16, 8, 8, 4, 2048, 702, 1, 20, 64, 8, 16) == bad
    assert lib.bslstm_key_switch_lwe(None, 4096, 1 << 20, 8, 8, 4, 1024, 630, 4, 5, 32,
                                     4096 + 8 * 1024, 1 << 21) == bad
    assert "overlap" in lib.bslstm_last_error().decode()
    assert lib.bslstm_key_switch_lwe(None, 4096, 1 << 20, 8, 8, 4, 1024, 630, 4, 5, 32, 1 << 22,
                                     1 << 20) == bad
    assert "overlap" in lib.bslstm_last_error().decode()
    assert _core.lib().bslstm_packing_workspace_bytes(2, 2048, 2048) == \
        2 * 2048 * 2048 * 8 + 2 * 256 * 2 * 1024 * 16


def test_split_form_entry_points_reject_on_host():
    """Validation of the split-form entry points runs before any GPU work."""
    lib = _core.lib()
    bad = _core.BSLSTM_INVALID_ARGUMENT
    dims = (_core._I64 * 6)(250, 2, 3, 2, 8, 1024)
    B = 16
    ws = lib.bslstm_mul_bootstrap_split_workspace_bytes(B, dims, 2048, 0)
    # 64 ciphertexts per workgroup at N = 2048: the smallest chunk
    assert ws == 64 * (64 * 2 * 48 * 64 + 2 * 64 * 64 * 8)
    assert lib.bslstm_mul_bootstrap_split_key_bytes(dims, 2048) == 250 * 64 * 48 * 2 * 5 * 128
    assert lib.bslstm_mul_bootstrap_split_key_bytes(dims, 4096) == -1

    def call(**kw):
        a = dict(ct_elems=4096, idx_elems=B * 250 * 8, batch=B, lb=7, lq=32, l=3, degree=2048,
                 ws_bytes=ws)
        a.update(kw)
        return lib.bslstm_mul_bootstrap_split(None, 8, a["ct_elems"], 8, dims, 8, a["idx_elems"], 8,
                                              a["batch"], a["lb"], a["lq"], a["l"], a["degree"], 8,
                                              a["ws_bytes"])
    assert call(lq=64) == bad and "log2_q <= 32" in lib.bslstm_last_error().decode()
    assert call(lb=8) == bad and "int8 digits" in lib.bslstm_last_error().decode()
    assert call(degree=4096) == bad and "Unsupported degree" in lib.bslstm_last_error().decode()
    assert call(l=2) == bad
    assert call(idx_elems=5) == bad and "indices" in lib.bslstm_last_error().decode()
    assert call(ws_bytes=ws - 1) == bad and "workspace" in lib.bslstm_last_error().decode()
    assert call(batch=0, idx_elems=0) == _core.BSLSTM_OK


def test_pair_form_entry_points_reject_on_host():
    """Validation of the pair-form entry points runs before any GPU work."""
    lib = _core.lib()
    bad = _core.BSLSTM_INVALID_ARGUMENT
    dims = (_core._I64 * 6)(250, 2, 3, 2, 8, 1024)
    B = 4
    # header (status word, flags, hello) padded to 256 bytes, then 8 pairs x
    # 2 step parities x 2 members x l x N/2 complex
    ws = lib.bslstm_mul_bootstrap_pair_workspace_bytes(B, 3, 2048)
    assert ws == 256 + 8 * 2 * 2 * 3 * 1024 * 16
    assert lib.bslstm_mul_bootstrap_pair_workspace_bytes(B, 4, 2048) == -1
    assert lib.bslstm_mul_bootstrap_pair_workspace_bytes(B, 3, 1024) == 256 + 8 * 2 * 2 * 3 * 512 * 16
    assert lib.bslstm_mul_bootstrap_pair_workspace_bytes(B, 3, 512) == -1

    def call(**kw):
        a = dict(ct_elems=4096, idx_elems=B * 250 * 8, batch=B, lb=7, lq=32, l=3, degree=2048,
                 ws_bytes=ws, spin=1000, dims=dims)
        a.update(kw)
        return lib.bslstm_mul_bootstrap_pair(None, 8, a["ct_elems"], 8, a["dims"], None, 0, 0, 8,
                                             a["idx_elems"], 8, a["batch"], a["lb"], a["lq"],
                                             a["l"], a["degree"], 8, a["ws_bytes"], a["spin"])
    assert call(degree=512, ct_elems=1024, dims=(_core._I64 * 6)(250, 2, 3, 2, 8, 256)) == bad
    assert "degree 1024 or 2048" in lib.bslstm_last_error().decode()
    assert call(l=4, dims=(_core._I64 * 6)(250, 2, 4, 2, 8, 1024)) == bad
    assert "l <= 3" in lib.bslstm_last_error().decode()
    assert call(spin=0) == bad and "spin_limit" in lib.bslstm_last_error().decode()
    assert call(spin=(1 << 30) + 1) == bad
    assert call(ws_bytes=ws - 1) == bad and "workspace" in lib.bslstm_last_error().decode()
    assert call(idx_elems=5) == bad and "indices" in lib.bslstm_last_error().decode()
    assert call(batch=0, idx_elems=0) == _core.BSLSTM_OK
    assert lib.bslstm_mul_bootstrap_pair_slots(None, 8, 4096, 8, dims, None, 2, 8, B * 250 * 8, 8,
                                               B, 7, 32, 3, 2048, 8, ws, 1000) == bad
