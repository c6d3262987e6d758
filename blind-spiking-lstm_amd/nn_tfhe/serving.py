"""Encrypted-inference serving with continuous batching across clients.

Requests are single encrypted sequences, each from one of many clients with
their own keys.  Every client's keys live in a slot of device-resident key
pools (bootstrapping key, int8 key-switching tiles, Galois keys, packing
key); a client is admitted when a slot is free or can be freed (no active
request uses it; least recently used first), so the HBM budget is the
number of slots times the per-client footprint.

The server advances in ticks.  At every tick the requests that have arrived
are admitted while request and key slots allow, and one LSTM timestep runs
for every active request at once: the bootstraps, key switches, trace
packings and packings of all clients share the same launches, each item
reading its client's key through its slot (the `_slots` entry points of the
library).  A request leaves when its last timestep is done.  Per-ciphertext
arithmetic does not depend on the batch it rides in, so a request's result
is bit-identical alone and in the mix.

This restates no reference code: the reference has no serving loop (its
inference scripts run one padded batch, run_speed.py:128-148).
"""
from __future__ import annotations

import time
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np
import torch

from cuTFHE import torus
from cuTFHE._core import TFHEInternalError
from cuTFHE.multiply import FORMS, PAIR_RUN_OUT, log2_int, mul_bootstrap, pair_error_word
from nn_tfhe.Spike_LSTM import CipherSpikeLSTM


class KeyPool:
    """Per-client keys in `slots` device-resident slots."""

    def __init__(self, slots: int, example_keys: dict, device="cuda"):
        self.device = torch.device(device)
        self.slots = slots
        self.pools = {k: torch.empty((slots,) + tuple(v.shape), dtype=v.dtype, device=self.device)
                      for k, v in example_keys.items()}
        self.slot_bytes = sum(v.numel() * v.element_size() for v in example_keys.values())
        self.owner: list = [None] * slots
        self.lru: OrderedDict = OrderedDict()        # client -> slot, least recent first
        self.pins = [0] * slots
        self.loads = 0

    def footprint_bytes(self) -> int:
        return self.slots * self.slot_bytes

    def acquire(self, client, key_provider):
        """Slot holding `client`'s keys (loaded on a miss), pinned; None if
        every slot is pinned by other clients."""
        if client in self.lru:
            slot = self.lru.pop(client)
            self.lru[client] = slot
        else:
            free = [i for i in range(self.slots) if self.owner[i] is None]
            if free:
                slot = free[0]
            else:
                victims = [c for c, s in self.lru.items() if self.pins[s] == 0]
                if not victims:
                    return None
                slot = self.lru.pop(victims[0])
            keys = key_provider(client)
            for k, pool in self.pools.items():
                pool[slot].copy_(keys[k])
            self.owner[slot] = client
            self.lru[client] = slot
            self.loads += 1
        self.pins[slot] += 1
        return slot

    def release(self, slot: int):
        self.pins[slot] -= 1


class SlotBootstrapper:
    """The bootstrapper interface CipherSpikeLSTM.step uses (prepare_LUT_sum,
    cuboot, packing, lut_fft), with every item's key read from a KeyPool
    slot: set_slots() gives the slot of each sequence of the next step."""

    def __init__(self, pool: KeyPool, dict_params, dict_params_lut, dict_params_ks_LWE,
                 dict_params_packing, collapse: int, n_lut: int):
        self.pool = pool
        self.device = pool.device
        self.dict_params = dict_params
        self.dict_params_lut = dict_params_lut
        self.dict_params_ks_LWE = dict_params_ks_LWE
        self.dict_params_packing = dict_params_packing
        self.collapse = collapse
        self.n_lut = n_lut
        self.seq_slots = None

    def set_slots(self, seq_slots):
        """Slots of the next step's sequences (a list or a tensor).  Host
        values go through a new pinned buffer with an asynchronous copy, so
        the host does not wait for the device's queued work (the caching host
        allocator keeps the buffer until the copy has run)."""
        if isinstance(seq_slots, torch.Tensor) and seq_slots.is_cuda:
            self.seq_slots = seq_slots.to(dtype=torch.int32)
            return
        host = torch.as_tensor(seq_slots, dtype=torch.int32).pin_memory()
        self.seq_slots = host.to(self.device, non_blocking=True)

    def _per(self, count: int) -> torch.Tensor:
        S = self.seq_slots.numel()
        if count % S:
            raise ValueError(f"{count} items for {S} sequences")
        return self.seq_slots.repeat_interleave(count // S)

    def lut_fft(self, lut) -> torch.Tensor:
        t = lut if isinstance(lut, torch.Tensor) else torch.as_tensor(np.asarray(lut))
        return torus.poly_fft(torus._i64(t.to(self.device)))

    def prepare_LUT_sum(self, c_lwe, lut_fft):
        p = self.dict_params_packing
        masks, bodies = c_lwe                                # (S, n_lut, G, N), (S, n_lut, G)
        cnt = masks.numel() // masks.shape[-1]
        rl = torus.pack_lwe_to_rlwe_slots(masks, bodies.reshape(masks.shape[:-1]),
                                          self.pool.pools["gk"], self._per(cnt),
                                          log2_int(p["beta_ks"]), p["l_ks"], log2_int(p["q"]))
        return torus.lut_product_sum(rl, lut_fft, masks.shape[-3], log2_int(p["q"]))

    def cuboot(self, h, encrypted_LUT):
        """Key switch + many-LUT bootstrap of rows ordered by sequence."""
        return self.cuboot_switched(self.key_switch(h), encrypted_LUT)

    def key_switch(self, h):
        """Key switch of rows ordered by sequence, each sequence's with its
        client's key.  Each sequence's rows are padded to a multiple of 32
        so that no 32-row tile mixes clients.  Returns (a (rows, n), b (rows, 1))."""
        pk, pb = h
        rows, n_in = pk.shape[0], pk.shape[-1]
        S = self.seq_slots.numel()
        per_seq = rows // S
        pad = -per_seq % 32
        a3, b2 = pk.reshape(S, per_seq, n_in), pb.reshape(S, per_seq)
        if pad:
            a3 = torch.cat([a3, torch.zeros(S, pad, n_in, dtype=a3.dtype, device=a3.device)], 1)
            b2 = torch.cat([b2, torch.zeros(S, pad, dtype=b2.dtype, device=b2.device)], 1)
        pk_ = self.dict_params_ks_LWE
        lq = log2_int(pk_["q"])
        a_ks, b_ks = torus.key_switch_mfma_slots(
            self.pool.pools["ks"], self.seq_slots.repeat_interleave((per_seq + pad) // 32),
            a3.reshape(-1, n_in), b2.reshape(-1), pk_["degree"], log2_int(pk_["beta_ks"]),
            pk_["l_ks"], lq)
        if pad:
            a_ks = a_ks.reshape(S, per_seq + pad, -1)[:, :per_seq].reshape(rows, -1)
            b_ks = b_ks.reshape(S, per_seq + pad)[:, :per_seq].reshape(rows)
        return a_ks, b_ks.reshape(rows, 1)

    def cuboot_switched(self, c_lwe_ks, encrypted_LUT):
        """cuboot after its key switch: many-LUT bootstrap of rows ordered by
        sequence, each with its client's bootstrapping key.  Bootstrap pairs
        share a key when a sequence has an even number of rows (else only the
        forms with one ciphertext per workgroup are tuned over)."""
        a_ks, b_ks = c_lwe_ks
        rows = a_ks.shape[0]
        per_seq = rows // self.seq_slots.numel()
        p = self.dict_params_lut
        idx = torus.rotation_indices(a_ks, log2_int(p["q"]), p["degree"], self.collapse, self.n_lut)
        lut = torch.stack([torus._i64(encrypted_LUT[0]), torus._i64(encrypted_LUT[1])], dim=1)
        br = mul_bootstrap(lut, self.pool.pools["bsk"], None, idx, log2_int(p["beta_bs"]),
                           log2_int(p["q"]), p["l_bs"], form="auto", slots=self._per(rows),
                           candidates=FORMS if per_seq % 2 == 0 else ("plain", "rows", "pair"))
        masks, bodies = torus.rotate_extract(br, b_ks.reshape(rows), log2_int(p["q"]), self.n_lut)
        return masks, bodies[..., None]

    def packing(self, x):
        p = self.dict_params
        masks, bodies = x
        lead = masks.shape[:-2]
        n_lwes, n_in = masks.shape[-2], masks.shape[-1]
        out = torus.packing_slots(masks.reshape(-1, n_lwes, n_in), bodies.reshape(-1, n_lwes),
                                  self.pool.pools["ksp"], self.seq_slots, log2_int(p["beta_ks"]),
                                  p["l_ks"], log2_int(p["q"]))
        return out[:, 0].reshape(*lead, -1), out[:, 1].reshape(*lead, -1)


@dataclass
class Request:
    rid: int
    client: int
    x: torch.Tensor                  # (T, 2, N) int64: the encrypted input frames
    arrival: float
    t: int = 0
    slot: int = -1
    H: torch.Tensor | None = None
    C: torch.Tensor | None = None
    L: torch.Tensor | None = None    # forget-gate LUT of the next timestep (G, 2, N)
    start: float = -1.0
    done: float = -1.0
    result: tuple = field(default_factory=tuple)


class EncryptedLSTMServer:
    """Continuous-batching server of one CipherSpikeLSTM layer."""

    def __init__(self, model_args: dict, pool: KeyPool, key_provider, max_active: int = 64):
        self.pool = pool
        self.key_provider = key_provider
        self.max_active = max_active
        a = model_args
        self.bs = SlotBootstrapper(pool, a["dict_params"], a["dict_params_lut"],
                                   a["dict_params_ks_LWE"], a["dict_params_packing"],
                                   a["dict_params_lut"]["collapse"], a["n_lut"])
        self.lstm = CipherSpikeLSTM(a["dict_params"], a["dict_params_lut"], a["input_dim"],
                                    a["hidden"], a["max_len"], a["linear_f"], a["linear_i"],
                                    a["linear_o"], self.bs, a["n_lut"], a["lut_f"], a["lut_i"],
                                    a["lut_o"], a["beta_x"])
        self.waiting: list = []
        self.active: list = []
        self.finished: list = []
        self.ticks = 0
        self.bootstraps = 0
        self.tick_events: list = []     # (start, end) HIP events of every tick
        # the pair-form error record as of the last tick, copied without a wait
        self._pair_host = torch.zeros(1, dtype=torch.int32, pin_memory=True)

    def submit(self, req: Request):
        self.waiting.append(req)

    def _admit(self, now: float):
        still = []
        for r in self.waiting:
            if r.arrival > now or len(self.active) >= self.max_active:
                still.append(r)
                continue
            slot = self.pool.acquire(r.client, self.key_provider)
            if slot is None:
                still.append(r)
                continue
            r.slot = slot
            r.start = now
            N = r.x.shape[-1]
            r.H = torch.zeros(2, N, dtype=torch.int64, device=self.pool.device)
            r.C = torch.zeros_like(r.H)
            self.active.append(r)
        self.waiting = still

    def tick(self, now: float) -> int:
        """Admit, then one timestep for every active request.  Returns the
        number of requests stepped."""
        self._admit(now)
        if not self.active:
            return 0
        act = self.active
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()
        H = torch.stack([r.H for r in act])
        C = torch.stack([r.C for r in act])
        X = torch.stack([r.x[r.t] for r in act])
        new = [i for i, r in enumerate(act) if r.L is None]
        if new:
            # first timestep of these requests: the forget-gate LUT of their
            # (zero) cell state, under each client's keys
            self.bs.set_slots([act[i].slot for i in new])
            Ln = self.lstm.forget_lut(torch.stack([act[i].C for i in new]))
            for j, i in enumerate(new):
                act[i].L = Ln[j]
        L = torch.stack([r.L for r in act])
        self.bs.set_slots([r.slot for r in act])
        # the step also prepares each request's next forget-gate LUT
        Hn, Cn, Ln = self.lstm.step_pipelined(H, C, X, L)
        self.ticks += 1
        self.bootstraps += len(act) * 3 * self.lstm.G
        keep = []
        for i, r in enumerate(act):
            r.H, r.C, r.L = Hn[i], Cn[i], Ln[i]
            r.t += 1
            if r.t == r.x.shape[0]:
                r.result = (r.H, r.C)
                r.L = None
                self.pool.release(r.slot)
                self.finished.append(r)
            else:
                keep.append(r)
        self.active = keep
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.tick_events.append((e0, e1))
        return len(act)

    def gpu_busy_s(self) -> float:
        """Device time of all ticks so far (waits for them)."""
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in self.tick_events) / 1e3

    def run(self, requests, clock=time.perf_counter):
        """Serve `requests` (arrival times in seconds from the start) in
        real time.  Returns the finished requests with start/done times.

        The host prepares and launches tick k+1 while the device runs tick k
        (all work is on one stream, so tick k+1 reads tick k's states, and a
        key slot freed at tick k is overwritten only after it); it waits for
        tick k before launching tick k+2.  A request's done time is when the
        host sees its last tick complete (at most one tick late)."""
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        t0 = clock()
        pending = None              # (event, requests whose last tick it ends) of the previous tick

        def settle():
            pending[0].synchronize()
            if int(self._pair_host[0]) & 1:
                # a blind rotation of some tick up to this one left wrong states
                raise TFHEInternalError(PAIR_RUN_OUT)
            end = clock() - t0
            for r in pending[1]:
                r.done = end

        while self.waiting or self.active:
            now = clock() - t0
            if not self.active and self.waiting and min(r.arrival for r in self.waiting) > now:
                if pending is not None:
                    settle()
                    pending = None
                    continue
                nxt = min(r.arrival for r in self.waiting)
                if clock() - t0 == now:                 # a clock that only ticks with work
                    now = nxt
                else:
                    time.sleep(min(0.01, nxt - now))
                    continue
            n_done = len(self.finished)
            stepped = self.tick(now)
            if stepped:
                word = pair_error_word(self.pool.device)
                if word is not None:
                    self._pair_host.copy_(word, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record()
                if pending is not None:
                    settle()
                pending = (ev, self.finished[n_done:])
        if pending is not None:
            settle()
        return self.finished
