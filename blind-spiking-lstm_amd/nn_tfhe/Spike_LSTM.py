"""Encrypted spiking LSTM of the reference's nn_tfhe/Spike_LSTM.py on the GPU
engine.

Same classes, constructor arguments and dataflow as the reference (Linear,
CipherSpikeLSTM, CipherMLP); ciphertexts are int64 torch tensors on the GPU
(torus classes mod q, the reference's q = 2^64 included) and every step runs
hand-written HIP kernels:

  Linear               bslstm_linear_extract (exact weight-polynomial product
                       and coefficient-0 extraction)
  hidden-state LWEs    bslstm_rotate_extract_i64 without rotation
  prepare_LUT (+ sum)  bslstm_pack_lwe_to_rlwe + bslstm_lut_product_sum
  cuboot               key switch + rotation exponents + blind rotation +
                       rotate/extract
  packing              bslstm_packing
  sums / rotations     bslstm_torus_axpy / bslstm_rotate_poly

The timestep additionally takes a batch of sequences (leading axis S): the
reference vmaps the whole model over sequences (run_speed.py:133); here one
set of launches serves all S sequences, so the blind rotations of a step run
as one batched launch.
"""
from __future__ import annotations

import numpy as np
import torch

from cuTFHE import torus
from cuTFHE.multiply import check_pair_errors, log2_int
from schemas.RLWE_jax import get_delta
from schemas.polynomial_jax import weights_to_polynomial


def _dev(x, device, dtype=None):
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
    return t.to(device=device, dtype=dtype) if dtype is not None else t.to(device)


class Linear:
    """nn_tfhe/Spike_LSTM.py:15-61: plaintext weights x RLWE, coefficient-0
    extraction, bias."""

    def __init__(self, output_dim: int, dict_params: dict, factor_W, device="cuda"):
        self.dict_params = dict_params
        self.delta = get_delta(dict_params)
        self.factor = factor_W
        self.degree = dict_params["degree"]
        self.log_q = log2_int(dict_params["q"])
        self.device = torch.device(device)
        self.kernel_polynomial = torch.zeros((output_dim, self.degree), dtype=torch.int64,
                                             device=self.device)
        self.bias = torch.zeros(output_dim, dtype=torch.int64, device=self.device)
        self.support = torch.zeros(0, dtype=torch.int32, device=self.device)

    def set_weights(self, W, B):
        """Spike_LSTM.py:27-41: W (in, out), B (out,) plaintext; each output row
        becomes round(factor_o * W[:, o]) as the polynomial
        w_0 - sum_{i>0} w_i X^(N-i), so that coefficient 0 of W_o(X) * m(X) is
        the dot product with the first `in` message coefficients."""
        W = np.asarray(W, np.float64)
        factor = np.broadcast_to(np.asarray(self.factor, np.float64), (W.shape[1],))
        Wt = np.rint(factor[:, None] * W.T).astype(np.int64)        # (out, in)
        out_dim, K = Wt.shape
        N = self.degree
        rows = np.stack([weights_to_polynomial(w) for w in Wt])      # (out, K)
        poly = np.zeros((out_dim, N), np.int64)
        poly[:, N - K + 1:] = rows[:, :-1]
        poly[:, 0] = rows[:, -1]
        self.kernel_polynomial = _dev(poly, self.device, torch.int64)
        self.support = torch.nonzero(self.kernel_polynomial.ne(0).any(dim=0)).reshape(-1).to(
            torch.int32)
        self.bias = _dev(np.rint(factor * np.asarray(B, np.float64)).astype(np.int64), self.device,
                         torch.int64)

    def __call__(self, x):
        """x = RLWE (A (..., N), B (..., N)) -> LWEs (masks (..., out, N),
        bodies (..., out, 1)), centred mod q (Spike_LSTM.py:46-61)."""
        A, B = x[0], x[1]
        lead = A.shape[:-1]
        ct = torch.stack([torus._i64(A).reshape(-1, self.degree),
                          torus._i64(B).reshape(-1, self.degree)], dim=1)
        masks, bodies = torus.linear_extract(ct, self.kernel_polynomial, self.bias, self.log_q,
                                             self.support, check_support=False)
        out = self.kernel_polynomial.shape[0]
        return masks.reshape(*lead, out, self.degree), bodies.reshape(*lead, out, 1)


class CipherSpikeLSTM:
    """nn_tfhe/Spike_LSTM.py:64-184: one encrypted spiking-LSTM layer.

    lut_f, lut_o: plaintext LUT polynomials (n_lut, hidden/n_lut, N) (0/+-1
    entries, setup_crypto.py:77,81); lut_i: [zeros, LUT values] trivial
    RLWE of the input gate (setup_crypto.py:78-80)."""

    def __init__(self, dict_params, dict_params_lut, input_dim, hidden_dim, max_len, linear_f,
                 linear_i, linear_o, heavyBS, n_lut, lut_f, lut_i, lut_o, beta_x=2 ** 53):
        self.dict_params = dict_params
        self.degree_lut = dict_params_lut["degree"]
        self.degree = dict_params["degree"]
        self.log_q = log2_int(dict_params["q"])
        self.input_dim = input_dim
        self.hidden_dim = hidden_dim
        self.max_len = max_len
        self.beta_x = beta_x
        self.W_f, self.W_i, self.W_o = linear_f, linear_i, linear_o
        self.heavyBS = heavyBS
        self.n_lut = n_lut
        dev = heavyBS.device
        G = hidden_dim // n_lut
        self.G = G
        # the LUT polynomials are plaintext constants: transform them once
        self.lut_f_fft = heavyBS.lut_fft(np.asarray(lut_f).reshape(n_lut, G, self.degree_lut))
        self.lut_o_fft = heavyBS.lut_fft(np.asarray(lut_o).reshape(n_lut, G, self.degree_lut))
        li = [torus._i64(_dev(np.rint(np.asarray(p, np.float64)), dev)) for p in lut_i]
        self.lut_i = torch.stack([li[0].reshape(G, -1), li[1].reshape(G, -1)], dim=1)  # (G, 2, N)
        # device-resident exponent of the H rotation (no host copy inside a step)
        self._shift = torch.full((1,), input_dim, dtype=torch.int64, device=dev)
        self._gates = None      # (weight tensors it was built from, merged weights)
        self._side = None       # side stream of step_pipelined (default priority)

    def _gate_weights(self):
        """The three gate Linears stacked into one weight set (rows f, i, o)
        with the union of their supports: one launch computes all three
        pre-activations (each row's result is the same as its own Linear's).
        Rebuilt when a Linear's weights are replaced (set_weights)."""
        lins = (self.W_f, self.W_i, self.W_o)
        key = tuple((id(l.kernel_polynomial), id(l.bias)) for l in lins)
        if self._gates is None or self._gates[0] != key:
            w = torch.cat([l.kernel_polynomial for l in lins])
            b = torch.cat([l.bias for l in lins])
            sup = torch.nonzero(w.ne(0).any(dim=0)).reshape(-1).to(torch.int32)
            self._gates = (key, (w, b, sup))
        return self._gates[1]

    def _lwe_add(self, x, y):
        return (torus.axpy(x[0], y[0], 1, 1, self.log_q), torus.axpy(x[1], y[1], 1, 1, self.log_q))

    def _spikes(self, masks, bodies, lo, hi):
        """boot rows lo:hi (G PBS x n_lut outputs) -> hidden LWEs ordered
        k*G + g (Spike_LSTM.py:141-148, 165-168)."""
        m = masks[:, lo:hi].permute(0, 2, 1, 3).reshape(masks.shape[0], -1, masks.shape[-1])
        b = bodies[:, lo:hi].permute(0, 2, 1).reshape(bodies.shape[0], -1)
        return m, b

    # The next step's forget-gate LUT goes to a side stream, overlapping the
    # output gate's bootstrap, when that bootstrap has at most this many
    # ciphertexts (its launch then leaves most of the chip idle).
    OVERLAP_MAX_BOOTSTRAPS = 128

    def forget_lut(self, C):
        """The forget gate's encrypted LUT from the cell state C (S, 2, N)
        (Spike_LSTM.py:113-132): the hidden-state LWEs of C (coefficients
        0..hidden-1, ordered k*G + g), trace-packed and multiplied by the
        plaintext LUTs, summed over the n_lut slots.  Returns (S, G, 2, N)."""
        S, N = C.shape[0], self.degree
        cm, cb = torus.rotate_extract(C, None, self.log_q, 1, self.hidden_dim)
        C_lwe = (cm.reshape(S, self.n_lut, self.G, N), cb.reshape(S, self.n_lut, self.G))
        return self.heavyBS.prepare_LUT_sum(C_lwe, self.lut_f_fft)

    def step(self, H, C, x_t):
        """One timestep (Spike_LSTM.py:108-178) for S sequences.
        H, C, x_t: RLWEs int64 (S, 2, N).  Returns (H_t, C_t) (S, 2, N)."""
        H_new, C_new, _ = self.step_pipelined(H, C, x_t, self.forget_lut(C), next_lut=False)
        return H_new, C_new

    def step_pipelined(self, H, C, x_t, lut_C, next_lut: bool = True):
        """The timestep of Spike_LSTM.py:108-178 as step(), with the forget
        gate's LUT given (forget_lut(C), or the third result of the previous
        step_pipelined call) and, if next_lut, the next step's LUT returned as
        well: (H_t, C_t, forget_lut(C_t)).
        That LUT depends only on the new cell state, so it is prepared while
        the output gate's bootstrap runs -- on a side stream when that
        bootstrap is small -- and leaves the sequential path of the next
        step.  The three gates share one key switch (exact, so batching it
        changes nothing).  Results are those of step(), bit for bit."""
        S, N = H.shape[0], self.degree
        G, nl, lq = self.G, self.n_lut, self.log_q
        # HX = x_t + H; the three gate pre-activations
        HX = torus.axpy(x_t, H, 1, 1, lq)
        w, bias, sup = self._gate_weights()
        gm, gb = torus.linear_extract(HX, w, bias, lq, sup, check_support=False)
        Gf, Gi = self.W_f.kernel_polynomial.shape[0], self.W_i.kernel_polynomial.shape[0]
        heavy = self.heavyBS
        # one key switch for the three gates (rows f, i, o of every sequence)
        ka, kb = heavy.key_switch((gm.reshape(-1, N), gb.reshape(-1, 1)))
        ka, kb = ka.reshape(S, -1, ka.shape[-1]), kb.reshape(S, -1, 1)
        fi_ks = (ka[:, :Gf + Gi].reshape(-1, ka.shape[-1]), kb[:, :Gf + Gi].reshape(-1, 1))
        o_ks = (ka[:, Gf + Gi:].reshape(-1, ka.shape[-1]), kb[:, Gf + Gi:].reshape(-1, 1))
        # first bootstrap: forget gate with the C-dependent LUT, input gate
        lut = torch.cat([lut_C, self.lut_i.unsqueeze(0).expand(S, G, 2, N)], dim=1)
        bm, bb = heavy.cuboot_switched(fi_ks, [lut[:, :, 0].reshape(-1, N),
                                               lut[:, :, 1].reshape(-1, N)])
        bm = bm.reshape(S, 2 * G, nl, N)
        bb = bb.reshape(S, 2 * G, nl)
        sI = self._spikes(bm, bb, G, 2 * G)
        sFC = self._spikes(bm, bb, 0, G)
        C_t = self._lwe_add(sFC, sI)                                            # (S, hidden, .)

        def pack_c():
            Cp = heavy.packing((C_t[0], C_t[1].unsqueeze(-1)))
            C_new = torch.stack([Cp[0], Cp[1]], dim=1)
            return C_new, (self.forget_lut(C_new) if next_lut else None)

        main = torch.cuda.current_stream()
        side = None
        if next_lut and S * G <= self.OVERLAP_MAX_BOOTSTRAPS:
            if self._side is None:
                self._side = torch.cuda.Stream(device=H.device)
            side = self._side
            side.wait_stream(main)
            with torch.cuda.stream(side):
                C_new, lut_next = pack_c()
        # second bootstrap: output gate with the C_t-dependent LUT
        lut_H = heavy.prepare_LUT_sum((C_t[0].reshape(S, nl, G, N),
                                       C_t[1].reshape(S, nl, G)), self.lut_o_fft)
        hm, hb = heavy.cuboot_switched(o_ks, [lut_H[:, :, 0].reshape(-1, N),
                                              lut_H[:, :, 1].reshape(-1, N)])
        H_t = self._spikes(hm.reshape(S, G, nl, N), hb.reshape(S, G, nl), 0, G)
        if side is None:
            C_new, lut_next = pack_c()
        # pack H back into an RLWE; it moves behind the input slots
        Hp = heavy.packing((H_t[0], H_t[1].unsqueeze(-1)))
        H_new = torus.rotate_poly(torch.stack([Hp[0], Hp[1]], dim=1), self._shift, lq)
        if side is not None:
            main.wait_stream(side)
            if not torch.cuda.is_current_stream_capturing():
                # made on the side stream, used on this one from here on
                for t_ in (C_new, lut_next):
                    t_.record_stream(main)
        return H_new, C_new, lut_next

    def __call__(self, x, seq_len=None):
        """Spike_LSTM.py:103-184: x = (A (T, N), B (T, N)) for one sequence or
        (S, T, N) for S sequences.  Returns ((H_T, C_T), (H_all, C_all)) as
        RLWE pairs, like the reference's scan."""
        A, B = torus._i64(x[0]), torus._i64(x[1])
        single = A.dim() == 2
        if single:
            A, B = A[None], B[None]
        S, T, N = A.shape
        H = torch.zeros(S, 2, N, dtype=torch.int64, device=A.device)
        C = torch.zeros_like(H)
        lut_C = self.forget_lut(C)
        Hs, Cs = [], []
        for t in range(T):
            x_t = torch.stack([A[:, t], B[:, t]], dim=1)
            H, C, lut_C = self.step_pipelined(H, C, x_t, lut_C, next_lut=t + 1 < T)
            Hs.append(H)
            Cs.append(C)
        Hs, Cs = torch.stack(Hs, dim=1), torch.stack(Cs, dim=1)                # (S, T, 2, N)
        check_pair_errors(A.device)       # wrong states from a run-out pair wait raise here

        def pair(t):
            return (t[..., 0, :], t[..., 1, :]) if not single else (t[0, ..., 0, :], t[0, ..., 1, :])
        return (pair(H), pair(C)), (pair(Hs), pair(Cs))


class CipherMLP:
    """nn_tfhe/Spike_LSTM.py:187-207: Linear -> ReLU bootstrap -> packing ->
    Linear."""

    def __init__(self, dict_params, dict_params_lut, hidden_dim, linear1, linear2, reluBS):
        self.dict_params = dict_params
        self.dict_params_lut = dict_params_lut
        self.hidden_dim = hidden_dim
        self.linear1 = linear1
        self.linear2 = linear2
        self.reluBS = reluBS

    def __call__(self, H_cipher):
        H1 = self.linear1(H_cipher)
        Y1 = self.reluBS(H1)
        H2 = self.reluBS.packing(Y1)
        return self.linear2(H2)
