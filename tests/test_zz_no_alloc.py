"""The allocation-free MFMA key switch (bslstm_key_switch_lwe_mfma_noalloc)
against the allocating entry point (bslstm_key_switch_lwe_mfma) and the exact
oracle: bit-identical outputs on two calls in a row, the second with a
smaller batch (the static scratch keeps the first call's digit tiles, which
must not leak into the second), and a clear InvalidArgument for a batch whose
digit tiles exceed the static array.  Collected last so that a failure here
cannot hide the other GPU tests.
"""
import numpy as np
import pytest
import torch

from oracle import lstm_oracle as L

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda:0")


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("log_q,lb,l,n_in,n_out,batches", [
    (64, 1, 20, 2048, 702, (300, 37)),     # LSTM key switch: 10 then 2 row tiles
    (32, 4, 5, 1024, 630, (256, 1)),       # 2048 -> 630-style config, one row at the end
])
def test_noalloc_key_switch_equals_allocating_form(cuda_device, log_q, lb, l, n_in, n_out, batches):
    from cuTFHE import _core, torus as T

    lib = _core.lib()
    cap = int(lib.bslstm_key_switch_lwe_mfma_noalloc_max_bytes())
    assert cap == 16 << 20
    rng = np.random.default_rng(n_in + l)
    s_old, s_new = rng.integers(0, 2, n_in), rng.integers(0, 2, n_out)
    ka, kb = L.keyswitch_key(rng, s_new, s_old, log_q, lb, l, 2.0 ** (log_q - 20))
    tiles = T.prepare_ks_key(_dev(ka), _dev(kb))
    for B in batches:
        assert int(lib.bslstm_ks_mfma_workspace_bytes(B, l, n_in)) <= cap
        a, b = L.lwe_encrypt(rng, rng.integers(-8, 8, B) << (log_q - 5), s_old, log_q,
                             2.0 ** (log_q - 30))
        da, db = _dev(a), _dev(b)
        ra, rb = T.key_switch_mfma(tiles, da, db, n_out, lb, l, log_q)
        ga = torch.full((B, n_out), -1, dtype=torch.int64, device="cuda:0")
        gb = torch.full((B,), -1, dtype=torch.int64, device="cuda:0")
        T.key_switch_mfma_noalloc(tiles, da, db, n_out, lb, l, log_q, ga, gb)
        np.testing.assert_array_equal(_host(ga), _host(ra))
        np.testing.assert_array_equal(_host(gb), _host(rb))
        wa, wb = L.key_switch_int(ka, kb, a, b, log_q, lb, l)
        np.testing.assert_array_equal(_host(ga), wa)
        np.testing.assert_array_equal(_host(gb), wb)


def test_noalloc_key_switch_rejects_batches_over_the_scratch(cuda_device):
    from cuTFHE import _core

    lib = _core.lib()
    cap = int(lib.bslstm_key_switch_lwe_mfma_noalloc_max_bytes())
    n_in, l = 2048, 20
    per_tile = int(lib.bslstm_ks_mfma_workspace_bytes(32, l, n_in))
    too_many = (cap // per_tile + 1) * 32
    x = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    st = lib.bslstm_key_switch_lwe_mfma_noalloc(
        _core.stream_of(x), x.data_ptr(), x.data_ptr(), x.data_ptr(), too_many, n_in, 702, 1, l,
        64, x.data_ptr(), x.data_ptr())
    assert st == _core.BSLSTM_INVALID_ARGUMENT
    assert "at most" in lib.bslstm_last_error().decode()
