"""On-disk client key files for the serving loop.

One file per client holds the four device-resident keys a KeyPool slot
needs (collapsed bootstrapping key, int8 key-switching tiles, Galois keys,
packing key), in safetensors format: plain tensors plus a string header, so
loading executes nothing from the file.  The header carries a format name
and version and the parameters the keys were made for; loading checks the
version, the parameters and every key's dtype and shape before the keys
reach the device.

The reference keeps keys in memory only and regenerates them from seeds
(setup_crypto.py:17-37); this file format has no counterpart there.
"""
from __future__ import annotations

import json
import math
import os

import torch

FORMAT = "bslstm-client-keys"
VERSION = 1
KEY_NAMES = ("bsk", "ks", "gk", "ksp")


class KeyFileError(ValueError):
    """A key file that is not of this format, version or parameter set."""


def _param_header(params) -> dict:
    dp, dpl, dpk, dpp = params
    return {"q_log2": int(round(math.log2(dp["q"]))), "N": int(dp["degree"]),
            "n": int(dpk["degree"]), "l_bs": int(dpl["l_bs"]),
            "beta_bs_log2": int(round(math.log2(dpl["beta_bs"]))),
            "collapse": int(dpl["collapse"]), "l_ks": int(dpk["l_ks"]),
            "beta_ks_log2": int(round(math.log2(dpk["beta_ks"]))),
            "l_gk": int(dpp["l_ks"]), "beta_gk_log2": int(round(math.log2(dpp["beta_ks"]))),
            "l_pk": int(dp["l_ks"]), "beta_pk_log2": int(round(math.log2(dp["beta_ks"])))}


def expected_shapes(params) -> dict:
    """(dtype, shape) of every key for a parameter set; the int8 key-switch
    tiles are checked by byte count (their tiling is the library's)."""
    h = _param_header(params)
    N, M = h["N"], h["N"] // 2
    m = 2 ** h["collapse"]
    return {"bsk": (torch.complex128, (h["n"] // h["collapse"], 2, h["l_bs"], 2, m, M)),
            "gk": (torch.complex128, (int(round(math.log2(N))), h["l_gk"], 2, M)),
            "ksp": (torch.complex128, (N, h["l_pk"], 2, M)),
            "ks": (torch.int8, None)}


def save_client_keys(path: str, keys: dict, params, client=None) -> None:
    """Writes one client's keys (tensors on any device) to `path`."""
    from safetensors.torch import save_file
    missing = [k for k in KEY_NAMES if k not in keys]
    if missing:
        raise KeyFileError(f"keys missing: {missing}")
    tensors, complex_keys = {}, []
    for name in KEY_NAMES:
        t = keys[name].detach().to("cpu").contiguous()
        if t.is_complex():
            complex_keys.append(name)
            t = torch.view_as_real(t).contiguous()
        tensors[name] = t
    header = {"format": FORMAT, "version": str(VERSION), "complex": json.dumps(complex_keys),
              "params": json.dumps(_param_header(params)),
              "client": "" if client is None else str(client)}
    tmp = path + ".tmp"
    save_file(tensors, tmp, metadata=header)
    os.replace(tmp, path)


def read_header(path: str) -> dict:
    from safetensors import safe_open
    with safe_open(path, framework="pt") as f:
        meta = f.metadata() or {}
    if meta.get("format") != FORMAT:
        raise KeyFileError(f"{path}: not a {FORMAT} file")
    if meta.get("version") != str(VERSION):
        raise KeyFileError(f"{path}: version {meta.get('version')}, this build reads {VERSION}")
    return meta


def load_client_keys(path: str, params, device="cuda", ks_bytes: int | None = None) -> dict:
    """Reads and checks one client's keys, returns them on `device`.
    ks_bytes: the expected size of the key-switching tiles, if known."""
    from safetensors.torch import load_file
    meta = read_header(path)
    want = _param_header(params)
    got = json.loads(meta["params"])
    if got != want:
        diff = {k: (got.get(k), v) for k, v in want.items() if got.get(k) != v}
        raise KeyFileError(f"{path}: keys made for other parameters (file, expected): {diff}")
    raw = load_file(path, device="cpu")
    complex_keys = set(json.loads(meta.get("complex", "[]")))
    shapes = expected_shapes(params)
    out = {}
    for name in KEY_NAMES:
        if name not in raw:
            raise KeyFileError(f"{path}: key '{name}' missing")
        t = raw[name]
        if name in complex_keys:
            if t.dtype != torch.float64 or t.shape[-1] != 2:
                raise KeyFileError(f"{path}: key '{name}' is not a complex pair array")
            t = torch.view_as_complex(t.contiguous())
        dtype, shape = shapes[name]
        if t.dtype != dtype or (shape is not None and tuple(t.shape) != shape):
            raise KeyFileError(f"{path}: key '{name}' is {t.dtype} {tuple(t.shape)}, "
                               f"expected {dtype} {shape}")
        if name == "ks" and ks_bytes is not None and t.numel() != ks_bytes:
            raise KeyFileError(f"{path}: key-switch tiles have {t.numel()} bytes, expected {ks_bytes}")
        out[name] = t.to(device)
    return out


class DiskKeyProvider:
    """key_provider for KeyPool/EncryptedLSTMServer: client c's keys are read
    from `<directory>/client_<c>.safetensors` when the pool loads them."""

    def __init__(self, directory: str, params, device="cuda", ks_bytes: int | None = None):
        self.directory = directory
        self.params = params
        self.device = device
        self.ks_bytes = ks_bytes
        self.reads = 0

    def path(self, client) -> str:
        return os.path.join(self.directory, f"client_{client}.safetensors")

    def __call__(self, client) -> dict:
        self.reads += 1
        return load_client_keys(self.path(client), self.params, self.device, self.ks_bytes)
