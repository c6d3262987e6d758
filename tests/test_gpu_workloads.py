"""The LSTM-path workloads at their real sizes, checked by decryption.

Reference parameters (parameters_crypto.py:6-33: q = 2^64, N = 2048,
n = 702, BSK l = 2 / Bg = 2^13, collapse 3, beta_x = 2^56), keys from
setup_crypto.sample_crypto_keys / sample_client_keys_device:
  * one spiking-LSTM layer, input 128, hidden 128, n_lut 4 (the one-layer
    workload), 3 timesteps, 1 and 4 sequences (the latter runs the
    one-group trace packing, >= 512 LWEs per launch);
  * the ReLU head (ReluBoostrapper, CipherMLP: Linear 256 -> 32, ramp LUT,
    packing, Linear 32 -> 2);
  * the classifier: 4 uneven-length sequences at hidden 256 through the
    timestep engine (captured steps), final states equal to each sequence
    run alone, then the head;
  * the continuous-batching server with 4 clients at N = 2048.
Every decrypted state is compared with the plaintext recurrence
(nn_tfhe/plaintext.py), evaluated on the decrypted previous states, for each
unit whose pre-activations lie at least MARGIN_SIGMAS standard deviations of
the LUT-index noise from every discontinuity of its LUTs; the rest may take
either neighbouring value.  That noise is the reference parameters' own
(lut_index_noise below): the modulus switch of the blind rotation rounds
each of the n/c group sums to the LUT grid, and the key switch adds
sqrt(N l_ks) * (beta_ks/2) * sigma_lwe.  At n_lut = 4 it is about 6.5 LUT
entries, at n_lut = 1 (the ReLU head) about 20 grid positions.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

S_LUT = 4.0            # LUT half-range (pre-activations in [-4, 4))
W_SCALE = 16           # pre-activation = (hx @ W) / 16 (Linear factor q/16/16/beta_x)
MOD = 256              # q / beta_x: values live mod 256


MARGIN_SIGMAS = 6.0


def lut_index_noise(n_lut: int) -> float:
    """LUT-index noise of the reference parameters (nn_tfhe/plaintext.py)."""
    import parameters_crypto as PC
    from nn_tfhe.plaintext import lut_index_noise as noise
    return noise(PC.dict_params_lut, PC.dict_params_ks_LWE, PC.degree, PC.q, n_lut)


def dev(x):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda:0")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def keys():
    import setup_crypto as SC
    return SC.sample_crypto_keys(11)


def _layer(keys, seed, input_dim, hidden, n_lut=4, max_len=64, s_head=None):
    import parameters_crypto as PC
    import setup_crypto as SC
    from nn_tfhe.Spike_LSTM import CipherSpikeLSTM
    from nn_tfhe.plaintext import PlainSpikingLSTM
    sk, ksk_packing, galois_key, bsk, ksk_bs, _ = keys
    G, K = hidden // n_lut, input_dim + hidden
    rng = np.random.default_rng(seed)
    heavy, relu = SC.get_tfhe_activation(bsk, ksk_bs, ksk_packing, galois_key,
                                         np.full(s_head or 1, S_LUT), n_lut, device="cuda:0")
    W = {c: rng.integers(-7, 8, (K, G)) for c in "fio"}          # int4 weights
    gamma = {c: rng.uniform(-2, 2, (n_lut, G)) for c in "fio"}
    eta = np.ones((n_lut, G), np.int64)
    factor = PC.q / W_SCALE / W_SCALE / PC.beta_x
    lin = [SC.make_linear(W[c], np.zeros(G), factor, PC.dict_params, device="cuda:0") for c in "fio"]
    sv = np.full(G, S_LUT)
    lut_f, lut_i, lut_o = SC.get_lut(n_lut, hidden, PC.degree_lut, sv, sv, sv, gamma["f"],
                                     gamma["i"], gamma["o"], eta)
    lstm = CipherSpikeLSTM(PC.dict_params, PC.dict_params_lut, input_dim, hidden, max_len, lin[0],
                           lin[1], lin[2], heavy, n_lut, lut_f, lut_i, lut_o, PC.beta_x)
    plain = PlainSpikingLSTM(W, {c: np.zeros(G) for c in "fio"}, gamma, eta, S_LUT, n_lut,
                             W_SCALE, MOD)
    entry = 2 * S_LUT / (PC.degree_lut // n_lut)                   # pre-activation per entry
    margin = MARGIN_SIGMAS * lut_index_noise(n_lut) * entry
    return lstm, plain, relu, margin


def _encrypt_spikes(sk, x, input_dim, rng):
    """Spike frames x (..., input_dim) -> RLWE (A, b) int64 (..., N)."""
    import parameters_crypto as PC
    from schemas.RLWE_jax import encrypt
    msg = np.zeros(x.shape[:-1] + (PC.degree,), np.int64)
    msg[..., :input_dim] = np.asarray(x, np.int64) << 56
    return encrypt(msg, sk, PC.dict_params, rng)


def _values(ct, sk, lo, hi):
    """Decrypted values (units of beta_x, mod 256) of coefficients lo:hi."""
    import parameters_crypto as PC
    from nn_tfhe.plaintext import wrap
    from schemas.RLWE_jax import _phase
    ph = _phase(ct, sk, PC.dict_params)
    return wrap(np.rint(np.asarray(ph)[..., lo:hi] / PC.beta_x).astype(np.int64), MOD)


def _check_steps(plain, margin, x, Hdec, Cdec):
    """x (T, ..., input_dim); Hdec/Cdec (T, ..., hidden) decrypted states.
    Each step's plaintext rule on the previous decrypted states; returns
    (checked, total)."""
    H = np.zeros_like(Hdec[0])
    C = np.zeros_like(Cdec[0])
    checked = total = 0
    for t in range(x.shape[0]):
        Hp, Cp, det = plain.step(x[t], H, C, margin)
        total += det.size
        checked += int(det.sum())
        bad = det & ((Cdec[t] != Cp) | (Hdec[t] != Hp))
        if bad.any():
            from nn_tfhe.plaintext import lut_margin
            pre = plain.pre(x[t], H)
            lines = []
            for idx in zip(*np.nonzero(bad)):
                u = idx[-1]
                k, g = divmod(int(u), plain.G)
                gi = idx[:-1] + (g,)
                lines.append(
                    f"step {t} unit {idx}: C {Cdec[t][idx]} vs {Cp[idx]}, H {Hdec[t][idx]} vs "
                    f"{Hp[idx]}, C_prev {C[idx]}; " + ", ".join(
                        f"{c}: pre {pre[c][gi]:.4f} gamma {plain.gamma[c][k, g]:.4f} margin "
                        f"{lut_margin(pre[c][gi], plain.gamma[c][k, g], plain.s):.4f}" for c in "fio"))
            raise AssertionError("\n".join(lines[:20]))
        H, C = Hdec[t], Cdec[t]
    return checked, total


@pytest.mark.parametrize("S", [1, 4])
def test_lstm_layer_hidden128_decrypts_to_the_plaintext_rule(cuda_device, keys, S):
    """One layer, input 128, hidden 128, n_lut 4, at the reference
    parameters: 3 timesteps of S sequences."""
    sk = keys[0]
    lstm, plain, _, margin = _layer(keys, 5, 128, 128)
    rng = np.random.default_rng(S)
    T_ = 3
    x = rng.integers(0, 2, (T_, S, 128))
    A, b = _encrypt_spikes(sk, x, 128, rng)
    xs = torch.stack([dev(A), dev(b)], dim=2)                    # (T, S, 2, N)
    H = torch.zeros(S, 2, lstm.degree, dtype=torch.int64, device="cuda:0")
    C = torch.zeros_like(H)
    lut = lstm.forget_lut(C)
    Hd, Cd = [], []
    for t in range(T_):
        H, C, lut = lstm.step_pipelined(H, C, xs[t], lut)
        Hd.append(_values((host(H[:, 0]), host(H[:, 1])), sk, 128, 256))
        Cd.append(_values((host(C[:, 0]), host(C[:, 1])), sk, 0, 128))
    checked, total = _check_steps(plain, margin, x, np.stack(Hd), np.stack(Cd))
    assert checked >= 0.3 * total, (checked, total)


def _head(relu, hidden, rng):
    """CipherMLP hidden -> 32 -> ReLU -> 2 with int8 weights scaled by 1/64
    and 1/128 (the ReLU inputs stay inside [-4, 4) and the logits inside the
    torus); returns (head, Wt1, Wt2) with the rounded effective weights."""
    import parameters_crypto as PC
    import setup_crypto as SC
    from nn_tfhe.Spike_LSTM import CipherMLP
    W1 = rng.integers(-127, 128, (hidden, 32))
    W2 = rng.integers(-127, 128, (32, 2))
    l1 = SC.make_linear(W1, np.zeros(32), 1 / 64, PC.dict_params, device="cuda:0")
    l2 = SC.make_linear(W2, np.zeros(2), 1 / 128, PC.dict_params, device="cuda:0")
    head = CipherMLP(PC.dict_params, PC.dict_params_lut, hidden, l1, l2, relu)
    return head, np.rint(W1 / 64).astype(np.int64), np.rint(W2 / 128).astype(np.int64)


def _head_check(head, Wt1, Wt2, Hc, Hvals, sk):
    """Run the head on RLWE Hc (values Hvals at coefficients 0..hidden-1);
    check the ReLU outputs against the ramp and the logits against the
    second Linear of the decrypted ReLU outputs."""
    import parameters_crypto as PC
    from nn_tfhe.plaintext import fold, relu_ramp
    from schemas.LWE_jax import lwe_phase
    H1 = head.linear1(Hc)
    Y1 = head.reluBS(H1)
    y = lwe_phase((host(Y1[0]), host(Y1[1])), sk, PC.dict_params) / PC.beta_x     # (..., 32)
    pre = (np.asarray(Hvals, np.int64) @ Wt1) / W_SCALE
    step = S_LUT / (PC.degree_lut // 2)                  # ramp value per LUT entry (n_lut 1)
    tol = MARGIN_SIGMAS * lut_index_noise(1) * step
    ok = np.abs(np.abs(fold(pre, S_LUT)) - S_LUT) >= tol
    want = relu_ramp(pre, S_LUT, PC.degree_lut // 2)
    assert ok.mean() >= 0.8, ok.mean()
    np.testing.assert_allclose(y[ok], want[ok], atol=tol)
    # the ramp's values, not noise: the typical error is far below the bound
    assert np.abs(y[ok] - want[ok]).mean() < tol / 3
    logits = head.linear2(head.reluBS.packing(Y1))
    lg = lwe_phase((host(logits[0]), host(logits[1])), sk, PC.dict_params) / PC.beta_x
    np.testing.assert_allclose(lg, y @ Wt2, atol=step)
    return lg


def test_relu_head_decrypts_to_the_ramp(cuda_device, keys):
    """ReluBoostrapper.__call__ and CipherMLP (activation.py:195-238,
    Spike_LSTM.py:187-207) on 3 encrypted hidden states of 256 units."""
    sk = keys[0]
    _, _, relu, _ = _layer(keys, 6, 128, 256, s_head=32)
    rng = np.random.default_rng(12)
    head, Wt1, Wt2 = _head(relu, 256, rng)
    Hv = rng.integers(0, 2, (3, 256))
    A, b = _encrypt_spikes(sk, Hv, 256, rng)
    _head_check(head, Wt1, Wt2, (dev(A), dev(b)), Hv, sk)


def test_classifier_uneven_lengths_engine_and_head(cuda_device, keys):
    """4 sequences of lengths 5, 2, 4, 3 at hidden 256: the engine's final
    states (steps replayed from captured graphs) equal each sequence run
    alone; every step of each sequence decrypts to the plaintext rule; the
    head's logits decrypt to the head applied to the decrypted final H."""
    from nn_tfhe.engine import SequenceEngine
    from schemas.text_jax import rotate_ciphertext
    sk = keys[0]
    lstm, plain, relu, margin = _layer(keys, 7, 128, 256, s_head=32)
    rng = np.random.default_rng(13)
    lens = (5, 2, 4, 3)
    xs, seqs = [], []
    for L_ in lens:
        x = rng.integers(0, 2, (L_, 128))
        A, b = _encrypt_spikes(sk, x, 128, rng)
        xs.append(x)
        seqs.append((dev(A), dev(b)))
    H, C = SequenceEngine(lstm, use_graph=True).run(seqs)
    Hh, Ch = host(H), host(C)
    checked = total = 0
    for i, (A, b) in enumerate(seqs):
        (H1, C1), (Hs, Cs) = lstm((A, b))
        np.testing.assert_array_equal(Hh[i], np.stack([host(H1[0]), host(H1[1])]))
        np.testing.assert_array_equal(Ch[i], np.stack([host(C1[0]), host(C1[1])]))
        Hd = _values((host(Hs[0]), host(Hs[1])), sk, 128, 384)
        Cd = _values((host(Cs[0]), host(Cs[1])), sk, 0, 256)
        c, t = _check_steps(plain, margin, xs[i], Hd, Cd)
        checked, total = checked + c, total + t
    assert checked >= 0.25 * total, (checked, total)
    head, Wt1, Wt2 = _head(relu, 256, np.random.default_rng(14))
    shift = torch.full((1,), -128, dtype=torch.int64, device="cuda:0")
    Hr = rotate_ciphertext((H[:, 0], H[:, 1]), shift)
    Hvals = _values((host(Hr[0]), host(Hr[1])), sk, 0, 256)
    _head_check(head, Wt1, Wt2, Hr, Hvals, sk)


def test_serving_at_n2048_alone_equals_mix(cuda_device):
    """Continuous batching at the reference parameters: 4 clients with their
    own device keys, 5 requests (one arriving mid-flight) through a key pool
    of 2 slots; each request's final states equal the request served alone,
    and its cell state decrypts under its own client's key to integers."""
    import parameters_crypto as PC
    import setup_crypto as SC
    from nn_tfhe.serving import EncryptedLSTMServer, KeyPool, Request
    rng = np.random.default_rng(21)
    nl, hid, inp = 4, 128, 128
    G = hid // nl
    keys, secrets = {}, {}
    for c in range(4):
        keys[c], secrets[c] = SC.sample_client_keys_device(300 + c, device="cuda:0")
    W = {c: rng.integers(-7, 8, (inp + hid, G)) for c in "fio"}
    gamma = {c: rng.uniform(-2, 2, (nl, G)) for c in "fio"}
    eta = np.ones((nl, G), np.int64)
    factor = PC.q / W_SCALE / W_SCALE / PC.beta_x
    lin = [SC.make_linear(W[c], np.zeros(G), factor, PC.dict_params, device="cuda:0") for c in "fio"]
    sv = np.full(G, S_LUT)
    lut_f, lut_i, lut_o = SC.get_lut(nl, hid, PC.degree_lut, sv, sv, sv, gamma["f"], gamma["i"],
                                     gamma["o"], eta)
    args = dict(dict_params=PC.dict_params, dict_params_lut=PC.dict_params_lut,
                dict_params_ks_LWE=PC.dict_params_ks_LWE, dict_params_packing=PC.dict_params_packing,
                n_lut=nl, input_dim=inp, hidden=hid, max_len=8, linear_f=lin[0], linear_i=lin[1],
                linear_o=lin[2], lut_f=lut_f, lut_i=lut_i, lut_o=lut_o, beta_x=PC.beta_x)
    specs = [(0, 3, 0.0), (1, 2, 0.0), (2, 4, 0.0), (3, 2, 0.0), (0, 2, 1.0)]
    cts = []
    for c, T_, _ in specs:
        x = rng.integers(0, 2, (T_, inp))
        A, b = _encrypt_spikes(secrets[c][0], x, inp, rng)
        cts.append(torch.stack([dev(A), dev(b)], dim=1))

    def serve(items, slots):
        pool = KeyPool(slots, keys[0], device="cuda:0")
        srv = EncryptedLSTMServer(args, pool, lambda c: keys[c], max_active=8)
        alone = len(items) == 1
        reqs = [Request(i, specs[i][0], cts[i], 0.0 if alone else specs[i][2]) for i in items]
        srv.run(reqs, clock=lambda: float(srv.ticks))
        return {r.rid: (host(r.result[0]), host(r.result[1])) for r in srv.finished}, pool

    mixed, pool = serve(range(len(specs)), slots=2)
    assert pool.loads >= 4
    for i in range(len(specs)):
        alone, _ = serve([i], slots=1)
        np.testing.assert_array_equal(mixed[i][0], alone[i][0], err_msg=f"H of request {i}")
        np.testing.assert_array_equal(mixed[i][1], alone[i][1], err_msg=f"C of request {i}")
        # the final cell state decrypts under its own client's key to
        # integer values (units of beta_x) no larger than the sequence length
        import parameters_crypto as PC
        from schemas.RLWE_jax import _phase
        sk = secrets[specs[i][0]][0]
        ph = _phase((mixed[i][1][0], mixed[i][1][1]), sk, PC.dict_params)[:hid] / PC.beta_x
        assert np.abs(ph - np.rint(ph)).max() < 0.05
        assert np.abs(np.rint(ph)).max() <= specs[i][1]
