"""Client key files (nn_tfhe/keystore.py): round trip, and every check that
keeps a wrong file away from the device.  CPU only."""
import json

import pytest
import torch

import parameters_crypto as PC
from nn_tfhe import keystore as K


def _params():
    return PC.make_params(degree=256, n_ks=48)


def _keys(params, seed=0):
    g = torch.Generator().manual_seed(seed)
    out = {}
    for name, (dtype, shape) in K.expected_shapes(params).items():
        if name == "ks":
            out[name] = torch.randint(-128, 128, (4096,), dtype=torch.int8, generator=g)
        else:
            out[name] = torch.complex(torch.randn(shape, dtype=torch.float64, generator=g),
                                      torch.randn(shape, dtype=torch.float64, generator=g))
    return out


def test_round_trip(tmp_path):
    params = _params()
    keys = _keys(params)
    path = str(tmp_path / "client_3.safetensors")
    K.save_client_keys(path, keys, params, client=3)
    got = K.load_client_keys(path, params, device="cpu", ks_bytes=4096)
    assert set(got) == set(K.KEY_NAMES)
    for name in K.KEY_NAMES:
        assert got[name].dtype == keys[name].dtype
        assert torch.equal(got[name], keys[name])
    meta = K.read_header(path)
    assert meta["client"] == "3" and json.loads(meta["params"])["N"] == 256


def test_provider_reads_per_client(tmp_path):
    params = _params()
    for c in (0, 1):
        K.save_client_keys(str(tmp_path / f"client_{c}.safetensors"), _keys(params, c), params, c)
    prov = K.DiskKeyProvider(str(tmp_path), params, device="cpu")
    a, b = prov(0), prov(1)
    assert prov.reads == 2 and not torch.equal(a["bsk"], b["bsk"])


def test_rejects_other_parameters(tmp_path):
    params = _params()
    path = str(tmp_path / "k.safetensors")
    K.save_client_keys(path, _keys(params), params)
    other = PC.make_params(degree=256, n_ks=48, l_bs=3)
    with pytest.raises(K.KeyFileError, match="other parameters"):
        K.load_client_keys(path, other, device="cpu")


def test_rejects_wrong_version_and_format(tmp_path):
    from safetensors.torch import save_file
    params = _params()
    path = str(tmp_path / "k.safetensors")
    K.save_client_keys(path, _keys(params), params)
    from safetensors import safe_open
    with safe_open(path, framework="pt") as f:
        meta = dict(f.metadata())
        tensors = {k: f.get_tensor(k) for k in f.keys()}
    meta["version"] = "99"
    save_file(tensors, path, metadata=meta)
    with pytest.raises(K.KeyFileError, match="version 99"):
        K.load_client_keys(path, params, device="cpu")
    save_file(tensors, path, metadata={"format": "something else"})
    with pytest.raises(K.KeyFileError, match="not a"):
        K.load_client_keys(path, params, device="cpu")


def test_rejects_wrong_shape_and_size(tmp_path):
    params = _params()
    keys = _keys(params)
    keys["gk"] = keys["gk"][:, :2]
    path = str(tmp_path / "k.safetensors")
    K.save_client_keys(path, keys, params)
    with pytest.raises(K.KeyFileError, match="'gk'"):
        K.load_client_keys(path, params, device="cpu")
    keys = _keys(params)
    K.save_client_keys(path, keys, params)
    with pytest.raises(K.KeyFileError, match="tiles"):
        K.load_client_keys(path, params, device="cpu", ks_bytes=1234)


def test_missing_key_is_refused_on_save(tmp_path):
    params = _params()
    keys = _keys(params)
    del keys["ksp"]
    with pytest.raises(K.KeyFileError, match="missing"):
        K.save_client_keys(str(tmp_path / "k.safetensors"), keys, params)
