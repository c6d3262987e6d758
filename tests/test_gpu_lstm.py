"""Parity of the LSTM-path HIP kernels (through the C ABI) with the CPU oracle
(oracle/lstm_oracle.py), and one encrypted spiking-LSTM timestep end to end.

Bars:
  * bit-exact: Linear + extraction, int64 key switch, int64 rotation
    indices, int64 rotate/extract, torus sums, X^e rotations (exact integer
    kernels);
  * FP64-transform kernels against the exact oracle: LUT products and
    packing (one decomposition of exact inputs) coefficient by coefficient
    within TOL_REL * q (the FP64 rounding of 64-bit torus values; a wrong
    index, sign or twiddle gives errors of order q), exactly equal at q = 2^32
    sizes where the transform is exact;
  * kernels that decompose their own FP64 results again (trace packing, the
    blind rotation) at q = 2^64: there the FP64 error (~2^30) exceeds half
    the last digit's step, so the next decomposition may pick the other digit
    at a rounding boundary and add a key row encrypting (almost) zero.  The
    ciphertext is then a different but equally valid encryption: the bar is
    the PHASE (b - <a, s>), within TOL_REL * q of the oracle's.  At q = 2^32
    the coefficients match exactly;
  * the timestep (reference structure and moduli, small ring): decrypted
    hidden and cell states equal the plaintext spiking-LSTM rule for every
    unit whose gate pre-activations lie clear of their thresholds.
"""
import numpy as np
import pytest
import torch

from oracle import exact as X
from oracle import lstm_oracle as L
from oracle import tfhe_oracle as O

pytestmark = pytest.mark.gpu

TOL_REL = 2.0 ** -20      # FP64 transform products of 64-bit torus values


def dev(x, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda:0")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def max_err(a, b, log_q):
    with np.errstate(over="ignore"):
        d = L.tmod(np.asarray(a, np.int64) - np.asarray(b, np.int64), log_q)
    return float(np.abs(d.astype(np.float64)).max(initial=0.0))


@pytest.fixture(scope="module")
def T():
    from cuTFHE import torus
    return torus


def test_poly_fft_matches_reference_transform(cuda_device, T):
    rng = np.random.default_rng(1)
    for N in (128, 256, 512, 1024, 2048, 4096):
        p = rng.integers(-2 ** 40, 2 ** 40, (3, N))
        got = host(T.poly_fft(dev(p)))
        want = O.negacyclic_fft(p.astype(np.float64))
        assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max(), N


@pytest.mark.parametrize("N,log_q,K,out,batch", [(256, 64, 40, 10, 3), (2048, 64, 384, 32, 2),
                                                 (2048, 32, 256, 17, 1), (512, 48, 512, 9, 2)])
def test_linear_extract_bit_exact(cuda_device, T, N, log_q, K, out, batch):
    rng = np.random.default_rng(N + K)
    W = rng.integers(-40, 41, (K, out))
    P = L.weights_to_polynomial(W, N)
    bias = rng.integers(-2 ** 20, 2 ** 20, out)
    ct = L.uniform(rng, (batch, 2, N), log_q)
    m, b = T.linear_extract(dev(ct), dev(P), dev(bias), log_q)
    m, b = host(m), host(b)
    for i in range(batch):
        wm, wb = L.linear_extract(ct[i], P, bias, log_q)
        np.testing.assert_array_equal(m[i], wm)
        np.testing.assert_array_equal(b[i], wb)


def test_linear_extract_broadcast_and_dense_support(cuda_device, T):
    rng = np.random.default_rng(7)
    N, log_q = 256, 64
    P = rng.integers(-5, 6, (4, N))                    # dense weight polynomials
    ct = L.uniform(rng, (2, N), log_q)
    m, b = T.linear_extract(dev(ct), dev(P), None, log_q)
    wm, wb = L.linear_extract(ct, P, None, log_q)
    np.testing.assert_array_equal(host(m), wm)
    np.testing.assert_array_equal(host(b), wb)


@pytest.mark.parametrize("log_q,lb,l,n_in,n_out,B", [(64, 1, 20, 2048, 702, 40), (64, 4, 8, 300, 97, 70),
                                                     (32, 4, 5, 1024, 630, 33)])
def test_key_switch_i64_bit_exact(cuda_device, T, log_q, lb, l, n_in, n_out, B):
    rng = np.random.default_rng(l * n_in)
    s_old, s_new = rng.integers(0, 2, n_in), rng.integers(0, 2, n_out)
    ka, kb = L.keyswitch_key(rng, s_new, s_old, log_q, lb, l, 2.0 ** (log_q - 20))
    a, b = L.lwe_encrypt(rng, rng.integers(-8, 8, B) << (log_q - 5), s_old, log_q, 2.0 ** (log_q - 30))
    ga, gb = T.key_switch(dev(ka), dev(kb), dev(a), dev(b), lb, l, log_q)
    wa, wb = L.key_switch_int(ka, kb, a, b, log_q, lb, l)
    np.testing.assert_array_equal(host(ga), wa)
    np.testing.assert_array_equal(host(gb), wb)


@pytest.mark.parametrize("log_q,lb,l,n_in,n_out,B", [(64, 1, 20, 2048, 702, 40), (64, 1, 20, 2048, 702, 700),
                                                     (64, 7, 9, 300, 97, 70), (32, 4, 5, 1024, 630, 33),
                                                     (64, 2, 3, 37, 5, 1), (64, 7, 9, 300, 97, 150),
                                                     (32, 4, 5, 1024, 630, 129)])
def test_key_switch_mfma_bit_exact(cuda_device, T, log_q, lb, l, n_in, n_out, B):
    """The int8 limb MFMA key switch equals the exact oracle (and so the
    VALU int64 key switch) bit for bit, including ragged tiles; B >= 33
    (two or more row tiles) runs the four-row-tile kernel, B = 40 and 70
    with idle waves, B = 150 with a lone last row tile."""
    rng = np.random.default_rng(l * n_in + B)
    s_old, s_new = rng.integers(0, 2, n_in), rng.integers(0, 2, n_out)
    ka, kb = L.keyswitch_key(rng, s_new, s_old, log_q, lb, l, 2.0 ** (log_q - 20))
    a, b = L.lwe_encrypt(rng, rng.integers(-8, 8, B) << (log_q - 5), s_old, log_q, 2.0 ** (log_q - 30))
    tiles = T.prepare_ks_key(dev(ka), dev(kb))
    ga, gb = T.key_switch_mfma(tiles, dev(a), dev(b), n_out, lb, l, log_q)
    wa, wb = L.key_switch_int(ka, kb, a, b, log_q, lb, l)
    np.testing.assert_array_equal(host(ga), wa)
    np.testing.assert_array_equal(host(gb), wb)


def test_rotation_indices_i64_bit_exact(cuda_device, T):
    rng = np.random.default_rng(3)
    for (N, n, c, n_lut, lq) in [(2048, 702, 3, 4, 64), (256, 48, 3, 4, 64), (1024, 630, 3, 1, 32)]:
        pk = L.uniform(rng, (50, n), lq)
        got = host(T.rotation_indices(dev(pk), lq, N, c, n_lut))
        np.testing.assert_array_equal(got, L.rotation_indices_int(pk, lq, N, c, n_lut))


def test_rotate_extract_i64_bit_exact(cuda_device, T):
    rng = np.random.default_rng(4)
    for (N, lq, n_lut, n_ext, with_body) in [(2048, 64, 4, 4, True), (256, 64, 4, 4, True),
                                             (2048, 64, 1, 128, False), (512, 32, 2, 2, True)]:
        br = L.uniform(rng, (9, 2, N), lq)
        body = L.uniform(rng, 9, lq) if with_body else None
        gm, gb = T.rotate_extract(dev(br), None if body is None else dev(body), lq, n_lut, n_ext)
        wm, wb = L.rotate_extract_int(br, body, lq, n_lut, n_ext)
        np.testing.assert_array_equal(host(gm), wm)
        np.testing.assert_array_equal(host(gb), wb)


def test_axpy_and_rotate_bit_exact(cuda_device, T):
    rng = np.random.default_rng(5)
    for lq in (64, 40):
        x, y = L.uniform(rng, (4, 2, 512), lq), L.uniform(rng, (4, 2, 512), lq)
        with np.errstate(over="ignore"):
            want = L.tmod(3 * x - 5 * y, lq)
        np.testing.assert_array_equal(host(T.axpy(dev(x), dev(y), 3, -5, lq)), want)
        e = np.array([0, 1, -7, 1500])
        got = host(T.rotate_poly(dev(x), dev(e), lq))
        for i in range(4):
            for c in range(2):
                np.testing.assert_array_equal(got[i, c], L.tmod(O.coef_rotation(x[i, c], e[i]), lq))


def _gkey_fft(T, gk):
    return T.poly_fft(dev(gk))                       # (logN, l, 2, N) -> (logN, l, 2, N/2)


@pytest.mark.parametrize("N,log_q,lb,l,sigma,cnt", [(256, 64, 2, 20, 2.0 ** 17, 5),
                                                    (2048, 64, 2, 20, 2.0 ** 17, 2),
                                                    (512, 32, 4, 8, 1.0, 4),
                                                    (128, 64, 8, 8, 2.0 ** 20, 3)])
def test_trace_pack_vs_exact(cuda_device, T, N, log_q, lb, l, sigma, cnt):
    rng = np.random.default_rng(N + l)
    S = rng.integers(0, 2, N)
    gk = L.galois_key(rng, S, log_q, lb, l, sigma)
    a, b = L.lwe_encrypt(rng, rng.integers(-8, 8, cnt) << (log_q - 6), S, log_q, sigma)
    got = host(T.pack_lwe_to_rlwe(dev(a), dev(b), _gkey_fft(T, gk), lb, l, log_q))
    for i in range(cnt):
        A, Bp = L.trace_pack(a[i], int(b[i]), gk, lb, l, log_q)
        if log_q <= 32:
            assert max_err(got[i, 0], A, log_q) == 0
            assert max_err(got[i, 1], Bp, log_q) == 0
        ph_got = L.rlwe_phase(got[i, 0], got[i, 1], S, log_q)
        ph_want = L.rlwe_phase(A, Bp, S, log_q)
        assert max_err(ph_got, ph_want, log_q) <= TOL_REL * 2.0 ** log_q


@pytest.mark.parametrize("N,log_q,lb,l", [(2048, 64, 2, 20), (1024, 64, 2, 20), (512, 32, 4, 5)])
def test_trace_pack_same_in_any_batch(cuda_device, T, N, log_q, lb, l):
    """Trace packing runs one thread group per LWE for 512 or more LWEs and
    two for fewer; an LWE's result must not depend on the batch: 640 LWEs
    in one call equal the same LWEs in calls of 128, bit for bit (l = 5
    has an odd number of level pairs)."""
    rng = np.random.default_rng(N * l)
    S = rng.integers(0, 2, N)
    gk = _gkey_fft(T, L.galois_key(rng, S, log_q, lb, l, 2.0 ** (log_q - 30)))
    a, b = dev(L.uniform(rng, (640, N), log_q)), dev(L.uniform(rng, 640, log_q))
    whole = host(T.pack_lwe_to_rlwe(a, b, gk, lb, l, log_q))
    parts = np.concatenate([host(T.pack_lwe_to_rlwe(a[i:i + 128], b[i:i + 128], gk, lb, l, log_q))
                            for i in range(0, 640, 128)])
    np.testing.assert_array_equal(whole, parts)


@pytest.mark.parametrize("N,log_q,S_,G,nl", [(2048, 64, 2, 8, 4), (256, 64, 1, 3, 4), (512, 32, 3, 2, 2)])
def test_lut_product_sum_vs_exact(cuda_device, T, N, log_q, S_, G, nl):
    rng = np.random.default_rng(N * G)
    rl = L.uniform(rng, (S_, nl, G, 2, N), log_q)
    lut = rng.integers(-1, 2, (nl, G, N))
    got = host(T.lut_product_sum(dev(rl), T.poly_fft(dev(lut)), nl, log_q))
    tol = 0 if log_q <= 32 else TOL_REL * 2.0 ** log_q
    for s in range(S_):
        want = L.lut_product_sum(rl[s].reshape(nl * G, 2, N), lut.reshape(nl * G, N), G, nl, log_q)
        assert max_err(got[s], want, log_q) <= tol


@pytest.mark.parametrize("N,n_in,n_lwes,log_q,lb,l,P", [(256, 256, 40, 64, 8, 4, 2),
                                                        (512, 512, 512, 64, 8, 4, 1),
                                                        (256, 200, 256, 32, 4, 4, 2)])
def test_packing_vs_exact(cuda_device, T, N, n_in, n_lwes, log_q, lb, l, P):
    rng = np.random.default_rng(N + n_lwes)
    S = rng.integers(0, 2, N)
    s = rng.integers(0, 2, n_in)
    ksk = L.packing_ksk(rng, s, S, log_q, lb, l, 2.0 ** 17 if log_q == 64 else 1.0)
    masks = L.uniform(rng, (P, n_lwes, n_in), log_q)
    bodies = L.uniform(rng, (P, n_lwes), log_q)
    kf = T.poly_fft(dev(ksk))                                 # (n_in, l, 2, N/2)
    got = host(T.packing(dev(masks), dev(bodies), kf, lb, l, log_q))
    # the transform-domain sum over n_in rows is rounded once; at q = 2^32 the
    # products are exact up to that single rounding of each coefficient
    tol = 2.0 if log_q <= 32 else TOL_REL * 2.0 ** log_q
    for p in range(P):
        A, Bp = L.packing(masks[p], bodies[p], ksk, lb, l, log_q)
        assert max_err(got[p, 0], A, log_q) <= tol
        assert max_err(got[p, 1], Bp, log_q) <= tol


def test_packing_decrypts_to_the_lwe_phases(cuda_device, T):
    rng = np.random.default_rng(9)
    N, log_q, lb, l = 2048, 64, 8, 4
    S = rng.integers(0, 2, N)
    ksk = L.packing_ksk(rng, S, S, log_q, lb, l, 2.0 ** 17)
    mu = rng.integers(-4, 4, 128) << 56
    a, b = L.lwe_encrypt(rng, mu, S, log_q, 2.0 ** 17)
    out = host(T.packing(dev(a[None]), dev(b[None]), T.poly_fft(dev(ksk)), lb, l, log_q))
    ph = L.rlwe_phase(out[0, 0], out[0, 1], S, log_q)
    assert np.array_equal(np.rint(ph[:128] / 2.0 ** 56).astype(np.int64), mu >> 56)
    assert np.abs(ph[128:].astype(np.float64)).max() < 2.0 ** 52


@pytest.mark.parametrize("form", ["plain", "batched", "rows"])
def test_blind_rotation_at_q64_vs_exact(cuda_device, form):
    """The int64 accumulator path at the reference's LSTM modulus 2^64
    (l = 2, beta = 2^13, c = 3) against the int128 oracle."""
    from cuTFHE import multiply
    rng = np.random.default_rng(21)
    N, n, Lv, lb, c, B = 512, 24, 2, 13, 3, 5
    s = rng.integers(0, 2, n)
    S = rng.integers(0, 2, N)
    bsk_coef = L.bsk_kernel_coef(rng, s, S, 64, lb, Lv, c, 2.0 ** 17)
    bsk_fft = O.negacyclic_fft(bsk_coef.astype(np.float64))
    ct = L.uniform(rng, (B, 2, N), 64)
    idx = rng.integers(0, 2 * N, (B, n // c, 2 ** c)).astype(np.int32)
    got = host(multiply.mul_bootstrap(dev(ct), dev(bsk_fft), None, dev(idx), lb, 64, Lv, form))
    want = X.blind_rotate(ct, bsk_coef, idx, lb, 64, Lv)
    for i in range(B):
        ph_got = L.rlwe_phase(got[i, 0], got[i, 1], S, 64)
        ph_want = L.rlwe_phase(want[i, 0], want[i, 1], S, 64)
        assert max_err(ph_got, ph_want, 64) <= TOL_REL * 2.0 ** 64


# ---------------------------------------------------------------------------
# one encrypted spiking-LSTM timestep
# ---------------------------------------------------------------------------

class _Model:
    """A small layer with the reference's structure and the reference's
    moduli (q = 2^64, beta_x = 2^56): N = 256, n = 48, n_lut = 4, 16 hidden
    units, 16 inputs.  Every LUT spans pre-activations [-4, 4) in 64 entries."""

    def __init__(self, seed=0):
        import parameters_crypto as PC
        self.N, self.n_ks, self.nl, self.hidden, self.input_dim = 256, 48, 4, 16, 16
        self.params = PC.make_params(degree=self.N, n_ks=self.n_ks, sigma_lwe=2.0 ** 30)
        self.dp, self.dpl, self.dpk, self.dpp = self.params
        self.beta_x = 2.0 ** 56
        self.s = 4.0
        self.G = self.hidden // self.nl
        rng = np.random.default_rng(seed)
        K = self.input_dim + self.hidden
        # weights in units of 1/16 of a pre-activation: x = (sum_i w_i*v_i + b)/16
        self.W = {g: rng.integers(-3, 4, (K, self.G)) for g in "fio"}
        self.b = {g: rng.integers(-8, 9, self.G) for g in "fio"}
        self.gamma = {g: rng.uniform(-2.0, 2.0, (self.nl, self.G)) for g in "fio"}
        self.eta = np.ones((self.nl, self.G))
        self.x = rng.integers(0, 2, (3, self.input_dim))                 # spikes, 3 steps

    def factor(self):
        """Puts sum_i w_i * (v_i * beta_x) at x * q/(4s) (x = sum w v / 16)."""
        return 2.0 ** 64 / (4 * self.s) / 16 / self.beta_x


def _build_device_layer(m: _Model):
    import setup_crypto as SC
    from nn_tfhe.Spike_LSTM import CipherSpikeLSTM
    sk, ksk_packing, galois_key, bsk, ksk_bs, sk_bsk = SC.sample_crypto_keys(7, m.params, m.beta_x)
    heavy, _ = SC.get_tfhe_activation(bsk, ksk_bs, ksk_packing, galois_key, np.ones(m.G), m.nl,
                                      params=m.params, device="cuda:0")
    lin = [SC.make_linear(m.W[g], m.b[g] * m.beta_x, m.factor(), m.dp, device="cuda:0")
           for g in "fio"]
    sv = np.full(m.G, m.s)
    lut_f, lut_i, lut_o = SC.get_lut(m.nl, m.hidden, m.N, sv, sv, sv, m.gamma["f"],
                                     m.gamma["i"], m.gamma["o"], m.eta, m.beta_x)
    lstm = CipherSpikeLSTM(m.dp, m.dpl, m.input_dim, m.hidden, 3, lin[0], lin[1], lin[2], heavy,
                           m.nl, lut_f, lut_i, lut_o, m.beta_x)
    return lstm, sk


def test_merged_gate_linear_equals_separate(cuda_device):
    """The step's one-launch gate Linear gives each gate exactly what its own
    Linear gives, and follows a set_weights replacement."""
    m = _Model()
    lstm, _ = _build_device_layer(m)
    rng = np.random.default_rng(9)
    hx = dev(rng.integers(-2 ** 63, 2 ** 63, size=(3, 2, m.N), dtype=np.int64))
    w, b, sup = lstm._gate_weights()
    from cuTFHE import torus
    gm, gb = torus.linear_extract(hx, w, b, 64, sup, check_support=False)
    for k, lin in enumerate((lstm.W_f, lstm.W_i, lstm.W_o)):
        sm, sb = lin((hx[:, 0], hx[:, 1]))
        assert torch.equal(gm[:, k * m.G:(k + 1) * m.G], sm)
        assert torch.equal(gb[:, k * m.G:(k + 1) * m.G], sb[..., 0])
    lstm.W_i.set_weights(np.zeros((m.input_dim + m.hidden, m.G)), np.ones(m.G))
    w2, b2, _ = lstm._gate_weights()
    assert not torch.equal(w2, w) and torch.equal(w2[m.G:2 * m.G], lstm.W_i.kernel_polynomial)


def test_lstm_timesteps_decrypt_to_the_plaintext_rule(cuda_device):
    """Three timesteps.  At each step the plaintext rule is evaluated on the
    decrypted previous states (so one unit called near a threshold does not
    poison later checks); every unit whose gate pre-activations lie at least
    4 LUT entries from their thresholds (and inside the LUT range) must
    decrypt exactly to it."""
    from schemas.RLWE_jax import _phase, encrypt
    m = _Model()
    lstm, sk = _build_device_layer(m)
    rng = np.random.default_rng(3)
    T_ = m.x.shape[0]
    msg = np.zeros((T_, m.N), np.int64)
    msg[:, :m.input_dim] = (m.x * (1 << 56)).astype(np.int64)
    A, b = encrypt(msg, sk, m.dp, rng)
    (_, _), (Hs, Cs) = lstm((dev(A), dev(b)), None)
    entry = 2 * m.s / (m.N // m.nl)
    H = np.zeros(m.hidden)
    C = np.zeros(m.hidden)
    checked = total = 0
    for t in range(T_):
        phC = _phase((host(Cs[0][t]), host(Cs[1][t])), sk, m.dp)
        phH = _phase((host(Hs[0][t]), host(Hs[1][t])), sk, m.dp)
        decC = np.rint(phC[:m.hidden] / 2.0 ** 56)
        decH = np.rint(phH[m.input_dim:m.input_dim + m.hidden] / 2.0 ** 56)
        assert np.abs(np.delete(phC, np.arange(m.hidden))).max() < 2.0 ** 54
        hx = np.concatenate([m.x[t], H])
        pre = {g: (hx @ m.W[g] + m.b[g]) / 16.0 for g in "fio"}
        for k in range(m.nl):
            for g in range(m.G):
                u = k * m.G + g
                total += 1
                xs = [pre[c][g] for c in "fio"]
                ds = [abs(pre[c][g] - m.gamma[c][k, g]) for c in "fio"]
                if max(abs(v) for v in xs) > m.s - 4 * entry or min(ds) < 4 * entry:
                    continue
                c_new = C[u] * (pre["f"][g] >= m.gamma["f"][k, g]) + \
                    m.eta[k, g] * (pre["i"][g] >= m.gamma["i"][k, g])
                h_new = c_new * (pre["o"][g] >= m.gamma["o"][k, g])
                assert decC[u] == c_new, (t, u, decC[u], c_new)
                assert decH[u] == h_new, (t, u, decH[u], h_new)
                checked += 1
        H, C = decH, decC
    assert checked >= total // 4, (checked, total)


@pytest.mark.parametrize("overlap", [True, False])
def test_pipelined_steps_equal_plain_steps(cuda_device, overlap):
    """step_pipelined (the forget gate's LUT prepared one step ahead, on a
    side stream when overlap is on) gives the states of plain step() calls
    bit for bit, for 3 sequences over 3 steps."""
    from schemas.RLWE_jax import encrypt
    m = _Model(seed=3)
    lstm, sk = _build_device_layer(m)
    lstm.OVERLAP_MAX_BOOTSTRAPS = 1 << 20 if overlap else 0
    rng = np.random.default_rng(9)
    S, T_ = 3, 3
    msg = np.zeros((T_ * S, m.N), np.int64)
    msg[:, :m.input_dim] = rng.integers(0, 2, (T_ * S, m.input_dim)) << 56
    A, b = encrypt(msg, sk, m.dp, rng)
    x = torch.stack([dev(A), dev(b)], dim=1).reshape(T_, S, 2, m.N)
    H = torch.zeros(S, 2, m.N, dtype=torch.int64, device="cuda:0")
    C = torch.zeros_like(H)
    H2, C2 = H.clone(), C.clone()
    lut = lstm.forget_lut(C2)
    for t in range(T_):
        H, C = lstm.step(H, C, x[t])
        H2, C2, lut = lstm.step_pipelined(H2, C2, x[t], lut)
        np.testing.assert_array_equal(host(H2), host(H), err_msg=f"H, step {t}")
        np.testing.assert_array_equal(host(C2), host(C), err_msg=f"C, step {t}")
        np.testing.assert_array_equal(host(lut), host(lstm.forget_lut(C)), err_msg=f"LUT, step {t}")


def test_engine_uneven_lengths_and_graph_replays_match_eager(cuda_device):
    """Sequences of lengths 3, 1, 2 through the timestep engine: each final
    state equals the state of that sequence run alone (eager), and replays of
    the captured step graphs (with eager work in between) reproduce the eager
    results bit for bit."""
    from nn_tfhe.engine import SequenceEngine
    from schemas.RLWE_jax import encrypt
    m = _Model(seed=1)
    lstm, sk = _build_device_layer(m)
    rng = np.random.default_rng(8)
    seqs = []
    for T_ in (3, 1, 2):
        msg = np.zeros((T_, m.N), np.int64)
        msg[:, :m.input_dim] = rng.integers(0, 2, (T_, m.input_dim)) << 56
        A, b = encrypt(msg, sk, m.dp, rng)
        seqs.append((dev(A), dev(b)))
    H0, C0 = SequenceEngine(lstm).run(seqs)
    H0, C0 = host(H0), host(C0)
    for i, (A, b) in enumerate(seqs):                       # each sequence alone
        (H1, C1), _ = lstm((A, b))
        np.testing.assert_array_equal(H0[i], np.stack([host(H1[0]), host(H1[1])]))
        np.testing.assert_array_equal(C0[i], np.stack([host(C1[0]), host(C1[1])]))
    eng = SequenceEngine(lstm, use_graph=True)
    for rep in range(3):
        Hg, Cg = eng.run(seqs)
        np.testing.assert_array_equal(host(Hg), H0, err_msg=f"graph replay {rep}")
        np.testing.assert_array_equal(host(Cg), C0, err_msg=f"graph replay {rep}")
        SequenceEngine(lstm).run(seqs[1:])                  # eager work between replays


def test_serving_results_same_alone_and_in_the_mix(cuda_device, tmp_path):
    """Continuous batching across clients: requests from three clients with
    their own device-generated keys (key slots), served together with
    admissions mid-flight and a key pool smaller than the client count (so a
    slot is evicted and reloaded), give bit-identical final states to each
    request served alone; every final cell state decrypts under its own
    client's key to small multiples of beta_x.  The mixed run reads the keys
    from client key files, the solo runs from memory."""
    import setup_crypto as SC
    from nn_tfhe.keystore import DiskKeyProvider, save_client_keys
    from nn_tfhe.serving import EncryptedLSTMServer, KeyPool, Request
    from schemas.RLWE_jax import _phase, encrypt
    m = _Model(seed=2)
    keys, secrets = {}, {}
    for c in range(3):
        keys[c], secrets[c] = SC.sample_client_keys_device(100 + c, m.params, device="cuda:0")
    sv = np.full(m.G, m.s)
    lut_f, lut_i, lut_o = SC.get_lut(m.nl, m.hidden, m.N, sv, sv, sv, m.gamma["f"], m.gamma["i"],
                                     m.gamma["o"], m.eta, m.beta_x)
    lin = [SC.make_linear(m.W[g], m.b[g] * m.beta_x, m.factor(), m.dp, device="cuda:0") for g in "fio"]
    args = dict(dict_params=m.dp, dict_params_lut=m.dpl, dict_params_ks_LWE=m.dpk,
                dict_params_packing=m.dpp, n_lut=m.nl, input_dim=m.input_dim, hidden=m.hidden,
                max_len=8, linear_f=lin[0], linear_i=lin[1], linear_o=lin[2], lut_f=lut_f,
                lut_i=lut_i, lut_o=lut_o, beta_x=m.beta_x)
    rng = np.random.default_rng(4)
    # (client, length, arrival tick): the last one arrives while others run
    specs = [(0, 3, 0.0), (1, 2, 0.0), (2, 4, 0.0), (0, 2, 1.0)]
    xs = []
    for c, T_, _ in specs:
        msg = np.zeros((T_, m.N), np.int64)
        msg[:, :m.input_dim] = rng.integers(0, 2, (T_, m.input_dim)) << 56
        A, b = encrypt(msg, secrets[c][0], m.dp, rng)
        xs.append(torch.stack([dev(A), dev(b)], dim=1))

    for c in range(3):
        save_client_keys(str(tmp_path / f"client_{c}.safetensors"), keys[c], m.params, c)
    disk = DiskKeyProvider(str(tmp_path), m.params, device="cuda:0",
                           ks_bytes=keys[0]["ks"].numel())

    def serve(items, slots, provider=lambda c: keys[c]):
        pool = KeyPool(slots, keys[0], device="cuda:0")
        srv = EncryptedLSTMServer(args, pool, provider, max_active=8)
        alone = len(items) == 1
        reqs = [Request(i, specs[i][0], xs[i], 0.0 if alone else specs[i][2]) for i in items]
        srv.run(reqs, clock=lambda: float(srv.ticks))           # time counted in ticks
        return {r.rid: (host(r.result[0]), host(r.result[1])) for r in srv.finished}, pool

    mixed, pool = serve(range(4), slots=2, provider=disk)
    assert pool.loads >= 3 and disk.reads == pool.loads   # three clients through two slots
    for i in range(4):
        alone, _ = serve([i], slots=1)
        np.testing.assert_array_equal(mixed[i][0], alone[i][0], err_msg=f"H of request {i}")
        np.testing.assert_array_equal(mixed[i][1], alone[i][1], err_msg=f"C of request {i}")
        ph = _phase((mixed[i][1][0], mixed[i][1][1]), secrets[specs[i][0]][0], m.dp)[:m.hidden]
        frac = np.abs(ph / 2.0 ** 56 - np.rint(ph / 2.0 ** 56))
        assert frac.max() < 0.05 and np.abs(np.rint(ph / 2.0 ** 56)).max() <= specs[i][1]


@pytest.mark.parametrize("N", [256, 1024, 2048])
def test_poly_ifft_inverts_poly_fft(cuda_device, T, N):
    """bslstm_poly_ifft(bslstm_poly_fft(p)) = p for integer polynomials whose
    transform is exact in FP64 (|p| < 2^40), and the device key generation's
    exact product through the pair equals the int128 schoolbook mod 2^64."""
    import setup_crypto as SC
    rng = np.random.default_rng(N)
    p = rng.integers(-2 ** 40, 2 ** 40, (5, N))
    np.testing.assert_array_equal(host(T.poly_ifft(T.poly_fft(dev(p)))), p)
    a = rng.integers(-2 ** 63, 2 ** 63, (3, N), dtype=np.int64)
    s = rng.integers(0, 2, N)
    got = host(SC._negacyclic_mul_key(dev(a), dev(s)))
    want = np.stack([X.negacyclic_mul(a[i], s, 64) for i in range(3)])
    np.testing.assert_array_equal(got, want)
