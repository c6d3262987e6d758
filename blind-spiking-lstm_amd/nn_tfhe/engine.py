"""Timestep engine for many encrypted sequences of uneven length.

The reference runs one jax.lax.scan per sequence (nn_tfhe/Spike_LSTM.py:182)
and vmaps it over a padded batch (run_speed.py:128-134), so every sequence
pays for the longest one.  Here sequences are ordered by length, and at
timestep t the live ones are a prefix of that order: one CipherSpikeLSTM.step
over the prefix puts all their bootstraps into the same batched launches,
and finished sequences drop out.  Each step can be replayed from a HIP graph
captured once per live-batch size (static input/state buffers, the step's
launches recorded on the capture stream), which removes the host launch
overhead of the ~40 launches per step.  Steps are chained through
CipherSpikeLSTM.step_pipelined: each step also prepares the next step's
forget-gate LUT (kept in a static buffer beside H and C), overlapped with
its own output-gate bootstrap.
"""
from __future__ import annotations

import torch

from cuTFHE import torus
from cuTFHE.multiply import check_pair_errors


class SequenceEngine:
    """Runs `lstm` (a CipherSpikeLSTM) over a set of encrypted sequences."""

    def __init__(self, lstm, use_graph: bool = False):
        self.lstm = lstm
        self.use_graph = use_graph
        self._graphs: dict = {}
        self._bufs = None
        self._pool = None      # one memory pool shared by the per-size graphs (replayed in turn)

    def _alloc(self, S, N, device):
        if self._bufs is None or self._bufs[0].shape[0] < S:
            self._graphs.clear()
            G = self.lstm.G
            self._bufs = tuple(torch.zeros(S, 2, N, dtype=torch.int64, device=device)
                               for _ in range(3)) + (      # H, C, x
                torch.zeros(S, G, 2, N, dtype=torch.int64, device=device),)   # forget-gate LUT
        return self._bufs

    def _step(self, s):
        H, C, X, L = self._bufs
        Hn, Cn, Ln = self.lstm.step_pipelined(H[:s], C[:s], X[:s], L[:s])
        H[:s].copy_(Hn)
        C[:s].copy_(Cn)
        L[:s].copy_(Ln)

    def _graph_for(self, s):
        if s not in self._graphs:
            H, C, X, L = self._bufs
            # eager warm-up at this size (lazy library state, launch-shape
            # tuning); the step leaves its inputs untouched
            st = torch.cuda.Stream()
            st.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(st):
                self.lstm.step_pipelined(H[:s], C[:s], X[:s], L[:s])
            torch.cuda.current_stream().wait_stream(st)
            if self._pool is None:
                self._pool = torch.cuda.graph_pool_handle()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, pool=self._pool):
                self._step(s)
            self._graphs[s] = g
        return self._graphs[s]

    def run(self, sequences, step_times: list | None = None):
        """sequences: list of RLWE sequences, each (A (T_i, N), B (T_i, N))
        int64 on the GPU.  Returns (H, C) int64 (S, 2, N): the final states
        in input order.  If step_times is a list, (live, ms) per step is
        appended (device time of the step's launches)."""
        S = len(sequences)
        lens = [int(a.shape[0]) for a, _ in sequences]
        order = sorted(range(S), key=lambda i: -lens[i])
        N = int(sequences[0][0].shape[-1])
        dev = sequences[0][0].device
        H, C, X, L = self._alloc(S, N, dev)
        H[:S].zero_()
        C[:S].zero_()
        L[:S].copy_(self.lstm.forget_lut(C[:S]))
        # inputs in length order, padded to the longest: (T_max, S, 2, N)
        T_max = max(lens)
        xs = torch.zeros(T_max, S, 2, N, dtype=torch.int64, device=dev)
        for rank, i in enumerate(order):
            a, b = sequences[i]
            xs[: lens[i], rank, 0] = torus._i64(a)
            xs[: lens[i], rank, 1] = torus._i64(b)
        sorted_lens = [lens[i] for i in order]
        stream = torch.cuda.current_stream()
        for t in range(T_max):
            live = sum(1 for L_ in sorted_lens if L_ > t)
            X[:live].copy_(xs[t, :live])
            if step_times is not None:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
            if self.use_graph:
                self._graph_for(live).replay()
            else:
                self._step(live)
            if step_times is not None:
                e1.record(stream)
                step_times.append((live, e0, e1))
        if step_times is not None:
            torch.cuda.synchronize()
            for k, rec in enumerate(step_times):
                if len(rec) == 3:
                    step_times[k] = (rec[0], rec[1].elapsed_time(rec[2]))
        inv = [0] * S
        for rank, i in enumerate(order):
            inv[i] = rank
        idx = torch.as_tensor(inv, device=dev)
        # a pair-form blind rotation whose partner wait ran out left wrong
        # states: fail here rather than return them (one synchronisation)
        check_pair_errors(dev)
        return H[:S].index_select(0, idx), C[:S].index_select(0, idx)
