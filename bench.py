#!/usr/bin/env python3
"""Batched programmable-bootstrap throughput on one MI355X.

Workload (one "step"): the many-LUT programmable bootstrap of 16384 LWE
ciphertexts -- rotation exponents, collapsed blind rotation, X^(-b) rotation
and extraction of 4 LWEs per ciphertext -- at n=750, k=1, N=2048, l=3,
log2 Bg=7, collapse 3, torus32, 4-bit signed messages, 4 LUTs (sign,
ReLU-clip, two steps) packed per test polynomial.  Keys and inputs are
synthetic (seeded) and resident in HBM before the timed region.

Prints one JSON line (see the repository's bench contract).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "blind-spiking-lstm_amd")

# workload parameters (batched PBS throughput configuration)
CFG = dict(B=16384, n=750, N=2048, L=3, log_beta=7, collapse=3, n_lut=4, msg_bits=4, log_q=32,
           sigma_lwe=2.0 ** 17, sigma_bsk=2.0 ** 5)

FP64_PEAK_TFLOPS = 78.6
INT8_PEAK_TOPS = 5000.0          # dense int8 MFMA

# HBM bytes per launch of the blind-rotation kernel at this workload (16384
# ciphertexts), from rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000), the
# scales read off a 1 GiB torch.neg calibration kernel in the same runs
# (batched: mean of both launches, plain: of the two after the first).
# None for other batch sizes.
HBM_TRAFFIC_BYTES = {"batched": 110.73e9 + 1.040e9, "plain": 202.07e9 + 0.537e9}
HBM_TRAFFIC_SOURCE = ("profiles/br_kernel_2048_3_pbs16384.md (rocprofv3 --pmc FETCH_SIZE x2 + "
                      "WRITE_SIZE x1, separate runs of bench.py at this workload)")


def fp64_flops_per_bootstrap(n, N, L, c):
    """Algorithmic FP64 work of one collapsed blind rotation in the folded
    N/2-point transform convention: (n/c)*[(2l+2)*5*(N/2)*log2(N/2) +
    16*l*N*(2^c+1)]."""
    M = N // 2
    return (n // c) * ((2 * L + 2) * 5 * M * math.log2(M) + 16 * L * N * (2 ** c + 1))


def split_int8_ops_per_bootstrap(n, N, L, c):
    """Int8 MFMA work of one split-form blind rotation: per step and bin
    (2R bins of Z[Y]/(Y^M+1), N = M*R) a (2M x P*M) Toeplitz-key by (P*M)
    digit GEMM column, P = 2l*2^c rotated digit polynomials, for 9 limb
    pairs; 2 operations per multiply-accumulate."""
    M, R = {512: (32, 16), 1024: (32, 32), 2048: (64, 32)}[N]
    P = 2 * L * 2 ** c
    return (n // c) * (2 * R) * (2 * M) * (P * M) * 9 * 2


MANY_LUT_FNS = (
    lambda x: 1 if x >= 0 else 0,          # sign
    lambda x: min(max(x, 0), 7),           # ReLU clipped to 3 bits
    lambda x: 1 if x >= -4 else 0,         # step at -4
    lambda x: 1 if x >= 4 else 0,          # step at +4
)


def setup(cfg, device):
    import numpy as np
    import torch
    sys.path.insert(0, PKG)
    from cuTFHE.multiply import prepare_bsk
    from nn_tfhe.utils_spike import many_lut_polynomial
    from schemas.RLWE_jax import bsk_to_fourier, get_boostrapping_key_collapse

    rng = np.random.default_rng(0)
    q = 2.0 ** cfg["log_q"]
    s = rng.integers(0, 2, cfg["n"])
    S = rng.choice([0.0, 1.0], size=cfg["N"])
    params = {"q": q, "beta_bs": 2 ** cfg["log_beta"], "l_bs": cfg["L"], "sigma": cfg["sigma_bsk"],
              "degree": cfg["N"]}
    BSK = get_boostrapping_key_collapse(S, s, params, rng, cfg["collapse"])
    bsk = prepare_bsk(torch.as_tensor(bsk_to_fourier(BSK)).to(device), cfg["n"], cfg["collapse"])
    del BSK
    B = cfg["B"]
    half = 1 << (cfg["log_q"] - 1)
    delta = 1 << (cfg["log_q"] - cfg["msg_bits"] - 1)
    msg = rng.integers(-(1 << (cfg["msg_bits"] - 1)), 1 << (cfg["msg_bits"] - 1), B)
    a = rng.integers(-half, half, size=(B, cfg["n"]), dtype=np.int64)
    e = np.rint(rng.standard_normal(B) * cfg["sigma_lwe"]).astype(np.int64)
    b = a @ s + msg * delta + e
    b = ((b + half) % (2 * half)) - half
    T = many_lut_polynomial(cfg["N"], MANY_LUT_FNS[: cfg["n_lut"]], cfg["msg_bits"], float(delta))
    lut = np.stack([np.zeros(cfg["N"]), T])
    inputs = {
        "pk": torch.as_tensor(a.astype(np.float64)).to(device),
        "b": torch.as_tensor(b.astype(np.float64)).reshape(B, 1).to(device),
        "lut": torch.as_tensor(np.broadcast_to(lut, (B, 2, cfg["N"])).copy()).to(torch.int64).to(device),
        "bsk": bsk,
    }
    host = {"S": S, "msg": msg, "delta": delta}
    return inputs, host


def run_gpu(args):
    import numpy as np
    import torch
    sys.path.insert(0, PKG)
    from cuTFHE.multiply import (mul_bootstrap, mul_bootstrap_split, prepare_split_key,
                                 rotation_indices)
    from schemas.RLWE_jax import rotate_extract

    cfg = dict(CFG)
    cfg["B"] = args.batch
    device = torch.device("cuda:0")
    inputs, host = setup(cfg, device)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    split_key = None
    if args.split:
        split_key = prepare_split_key(inputs["bsk"], cfg["L"], cfg["N"])

    def step(form, br_events=None):
        idx = rotation_indices(inputs["pk"], cfg["log_q"], cfg["N"], cfg["collapse"], cfg["n_lut"])
        if br_events is not None:
            br_events[0].record(stream)
        if form == "split":
            br = mul_bootstrap_split(inputs["lut"], split_key, inputs["bsk"].shape, idx,
                                     cfg["log_beta"], cfg["log_q"], cfg["L"])
        else:
            br = mul_bootstrap(inputs["lut"], inputs["bsk"], None, idx, cfg["log_beta"],
                               cfg["log_q"], cfg["L"], form=form)
        if br_events is not None:
            br_events[1].record(stream)
        return rotate_extract(br, inputs["b"], cfg["log_q"], cfg["n_lut"])

    def timed(form, steps, warmup):
        for _ in range(warmup):
            step(form)
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            out = step(form, evs[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        br_ms = [a.elapsed_time(b) for a, b in evs]
        return out, (t1 - t0) * 1e3 / steps, float(np.mean(br_ms))

    B = cfg["B"]
    forms = {}
    if args.form in ("plain", "both"):
        _, ms_plain, br_plain = timed("plain", max(1, args.steps // 2) if args.form == "both" else
                                      args.steps, 1 if args.form == "both" else args.warmup)
        forms["plain_fp64"] = {"bootstraps_per_s": round(B / (ms_plain / 1e3), 2),
                               "ms_per_step": round(ms_plain, 3), "kernel_avg_ms": round(br_plain, 3)}
    main_form = "plain" if args.form == "plain" else "batched"
    if main_form == "batched":
        out, ms_per_step, br_avg_ms = timed("batched", args.steps, args.warmup)
        forms["batched_fp64"] = {"bootstraps_per_s": round(B / (ms_per_step / 1e3), 2),
                                 "ms_per_step": round(ms_per_step, 3),
                                 "kernel_avg_ms": round(br_avg_ms, 3)}
    else:
        out, ms_per_step, br_avg_ms = step("plain"), ms_plain, br_plain
    value = B / (ms_per_step / 1e3)
    if args.split:
        sout, ms_split, br_split = timed("split", 1, 1)
        ops = split_int8_ops_per_bootstrap(cfg["n"], cfg["N"], cfg["L"], cfg["collapse"]) * B
        ach = ops / (br_split / 1e3) / 1e12
        forms["split_int8"] = {
            "bootstraps_per_s": round(B / (ms_split / 1e3), 2), "ms_per_step": round(ms_split, 3),
            "form_avg_ms": round(br_split, 3),
            "equals_batched_fp64": bool(torch.equal(sout[0], out[0]) and torch.equal(sout[1], out[1])),
            "roofline": {"bound": "int8-mfma", "achieved": round(ach, 1), "peak": INT8_PEAK_TOPS,
                         "unit": "TOPS", "frac": round(ach / INT8_PEAK_TOPS, 4),
                         "note": "algorithmic int8 MFMA operations over the whole split form "
                                 "(GEMM and transform kernels)"}}
        if "plain_fp64" in forms:
            forms["split_over_plain"] = round(forms["split_int8"]["bootstraps_per_s"] /
                                              forms["plain_fp64"]["bootstraps_per_s"], 3)
        if "batched_fp64" in forms and "plain_fp64" in forms:
            forms["batched_over_plain"] = round(forms["batched_fp64"]["bootstraps_per_s"] /
                                                forms["plain_fp64"]["bootstraps_per_s"], 3)
        split_key = None

    # decrypt a sample of the outputs on the host (sanity, not the parity test)
    masks, bodies = out
    ns = min(B, 512)
    mk = masks[:ns].cpu().numpy().astype(np.int64)
    bd = bodies[:ns].cpu().numpy().astype(np.int64)
    ph = bd - mk @ host["S"].astype(np.int64)
    ph = ((ph + 2 ** 31) % 2 ** 32) - 2 ** 31
    dec = np.rint(ph / host["delta"]).astype(np.int64)
    exp = np.array([[f(int(x)) for f in MANY_LUT_FNS[: cfg["n_lut"]]] for x in host["msg"][:ns]])
    ok_frac = float((np.mod(dec, 32) == np.mod(exp, 32)).mean())

    flops = fp64_flops_per_bootstrap(cfg["n"], cfg["N"], cfg["L"], cfg["collapse"]) * B
    achieved = flops / (br_avg_ms / 1e3) / 1e12
    res = {
        "metric": "bootstraps/s @N=2048 (split vs plain FP64-FFT)",
        "value": round(value, 2),
        "unit": "bootstraps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded keys and ciphertexts)",
        "config": {"workload": "batched many-LUT PBS (value: the fastest form, batched FP64)",
                   "batch": B, "n": cfg["n"], "N": cfg["N"],
                   "l": cfg["L"], "log2_Bg": cfg["log_beta"], "collapse": cfg["collapse"],
                   "n_lut": cfg["n_lut"], "torus": cfg["log_q"], "form": main_form + "_fp64"},
        "roofline": {"bound": "fp64-valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": (HBM_TRAFFIC_BYTES[main_form] if B == CFG["B"] else None),
                     "traffic_source": HBM_TRAFFIC_SOURCE,
                     "kernel": f"br_kernel<2048,3,{1 if main_form == 'plain' else 2},1,1,8>",
                     "kernel_avg_ms": round(br_avg_ms, 3)},
        "decrypt_ok_sample": ok_frac,
        "forms": forms,
    }
    return res


def run_form_sweep():
    """Blind-rotation time of the launch forms at several batch sizes (the pair
    form where its co-residency limit admits the batch), at
    the batched-PBS parameters (torus32, l=3) and at the LSTM's (q=2^64, l=2,
    Bg=2^13, n=702): the data the autotuner's choices rest on.  Random keys and
    inputs (the time does not depend on their values); 3-launch minimum."""
    import torch
    sys.path.insert(0, PKG)
    from cuTFHE.multiply import check_pair_errors, mul_bootstrap, pair_applicable
    dev = torch.device("cuda:0")
    N = 2048
    out = {}
    for name, n_iter, L, lb, lq, batches in (("torus32 l=3 n=750", 250, 3, 7, 32, (64, 256, 1024)),
                                             ("q=2^64 l=2 n=702", 234, 2, 13, 64, (32, 64, 128, 1024))):
        g = torch.Generator().manual_seed(0)
        bsk = (torch.randn(n_iter, 2, L, 2, 8, N // 2, dtype=torch.complex128, generator=g)
               * 2.0 ** 30).to(dev)
        rows = {}
        for B in batches:
            ct = torch.randint(-2 ** 31, 2 ** 31, (B, 2, N), dtype=torch.int64, generator=g).to(dev)
            idx = torch.randint(0, 2 * N, (B, n_iter, 8), dtype=torch.int32, generator=g).to(dev)
            ms = {}
            forms = ("plain", "batched", "rows") + \
                (("pair",) if pair_applicable(N, L, B, dev) else ())
            for form in forms:
                mul_bootstrap(ct, bsk, None, idx, lb, lq, L, form=form)
                ts = []
                for _ in range(3):
                    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    a.record()
                    mul_bootstrap(ct, bsk, None, idx, lb, lq, L, form=form)
                    b.record()
                    b.synchronize()
                    ts.append(a.elapsed_time(b))
                ms[form] = round(min(ts), 3)
            rows[str(B)] = ms
        check_pair_errors()
        out[name] = rows
        del bsk
    return {"unit": "ms per launch", "by_batch": out}


# one encrypted spiking-LSTM layer at the reference's parameters
# (parameters_crypto.py: q = 2^64, N = 2048, n = 702, l_bs = 2, Bg = 2^13,
# c = 3; key switch 2^1 x 20; packing 2^8 x 4; trace packing 2^2 x 20)
# ---------------------------------------------------------------------------
# LSTM-path helpers: the synthetic layer of the one-layer workload and
# decryption checks against the plaintext recurrence (nn_tfhe/plaintext.py)
# ---------------------------------------------------------------------------
LSTM_S = 4.0          # LUT half-range: pre-activations in [-4, 4)
LSTM_W_SCALE = 16     # pre-activation = (hx @ W) / 16
MARGIN_SIGMAS = 6.0   # determined units: this many LUT-index noise std from any discontinuity


def synthetic_layer(rng, heavy, input_dim, hidden, n_lut, max_len):
    """int4 gate weights U[-7, 7] scaled so that a unit weight moves a
    pre-activation by 1/16 (the role of beta_w/s in setup_crypto.py:39-67),
    thresholds U[-2, 2], eta = 1.  Returns (CipherSpikeLSTM, plaintext model)."""
    import numpy as np
    import parameters_crypto as PC
    import setup_crypto as SC
    from nn_tfhe.Spike_LSTM import CipherSpikeLSTM
    from nn_tfhe.plaintext import PlainSpikingLSTM
    G, K = hidden // n_lut, input_dim + hidden
    factor = PC.q / LSTM_W_SCALE / LSTM_W_SCALE / PC.beta_x
    W = {c: rng.integers(-7, 8, (K, G)) for c in "fio"}
    lin = [SC.make_linear(W[c], np.zeros(G), factor, PC.dict_params, device=heavy.device)
           for c in "fio"]
    gam = {c: rng.uniform(-2, 2, (n_lut, G)) for c in "fio"}
    eta = np.ones((n_lut, G), np.int64)
    sv = np.full(G, LSTM_S)
    lut_f, lut_i, lut_o = SC.get_lut(n_lut, hidden, PC.degree_lut, sv, sv, sv, gam["f"], gam["i"],
                                     gam["o"], eta)
    lstm = CipherSpikeLSTM(PC.dict_params, PC.dict_params_lut, input_dim, hidden, max_len, lin[0],
                           lin[1], lin[2], heavy, n_lut, lut_f, lut_i, lut_o, PC.beta_x)
    plain = PlainSpikingLSTM(W, {c: np.zeros(G) for c in "fio"}, gam, eta, LSTM_S, n_lut,
                             LSTM_W_SCALE, int(PC.q / PC.beta_x))
    return lstm, plain


def lstm_margin(n_lut):
    import parameters_crypto as PC
    from nn_tfhe.plaintext import lut_index_noise
    sigma = lut_index_noise(PC.dict_params_lut, PC.dict_params_ks_LWE, PC.degree, PC.q, n_lut)
    return MARGIN_SIGMAS * sigma * 2 * LSTM_S / (PC.degree_lut // n_lut), sigma


def decrypt_values(ct, sk, lo, hi):
    """Decrypted values (units of beta_x, mod q/beta_x) of coefficients lo:hi."""
    import numpy as np
    import parameters_crypto as PC
    from nn_tfhe.plaintext import wrap
    from schemas.RLWE_jax import _phase
    ph = _phase(ct, sk, PC.dict_params)
    return wrap(np.rint(np.asarray(ph)[..., lo:hi] / PC.beta_x).astype(np.int64),
                int(PC.q / PC.beta_x))


def count_determined(plain, margin, x, Hdec, Cdec):
    """Per step, the plaintext rule on the previous decrypted states:
    (determined units, of those decrypting to the rule, all units)."""
    import numpy as np
    H, C = np.zeros_like(Hdec[0]), np.zeros_like(Cdec[0])
    det_n = ok_n = tot = 0
    for t in range(x.shape[0]):
        Hp, Cp, det = plain.step(x[t], H, C, margin)
        det_n += int(det.sum())
        ok_n += int((det & (Hdec[t] == Hp) & (Cdec[t] == Cp)).sum())
        tot += det.size
        H, C = Hdec[t], Cdec[t]
    return det_n, ok_n, tot


def lstm_decrypt_check(lstm, plain, sk, S, steps, rng):
    """`steps` timesteps of S sequences with fresh spike inputs; every step's
    states decrypted and compared with the plaintext rule."""
    import numpy as np
    import torch
    import parameters_crypto as PC
    from schemas.RLWE_jax import encrypt
    inp, hid, N = lstm.input_dim, lstm.hidden_dim, lstm.degree
    x = rng.integers(0, 2, (steps, S, inp))
    msg = np.zeros((steps, S, N), np.int64)
    msg[..., :inp] = x.astype(np.int64) << 56
    A, b = encrypt(msg, sk, PC.dict_params, rng)
    dev = lstm.heavyBS.device
    xs = torch.stack([torch.as_tensor(A), torch.as_tensor(b)], dim=2).to(dev)
    H = torch.zeros(S, 2, N, dtype=torch.int64, device=dev)
    C = torch.zeros_like(H)
    lut = lstm.forget_lut(C)
    Hd, Cd = [], []
    for t in range(steps):
        H, C, lut = lstm.step_pipelined(H, C, xs[t], lut)
        Hd.append(decrypt_values((H[:, 0].cpu().numpy(), H[:, 1].cpu().numpy()), sk, inp, inp + hid))
        Cd.append(decrypt_values((C[:, 0].cpu().numpy(), C[:, 1].cpu().numpy()), sk, 0, hid))
    margin, sigma = lstm_margin(lstm.n_lut)
    det, ok, tot = count_determined(plain, margin, x, np.stack(Hd), np.stack(Cd))
    return {"sequences": S, "steps": steps, "units": tot, "determined_frac": round(det / tot, 4),
            "determined_correct_frac": round(ok / max(det, 1), 6),
            "lut_index_noise_entries": round(sigma, 2), "margin_sigmas": MARGIN_SIGMAS}


LSTM_CFG = dict(input_dim=128, hidden=128, n_lut=4, seed=0)


def run_lstm(args):
    """Per-timestep latency of the encrypted LSTM step for `batch` sequences,
    steps chained as a sequence runs them (CipherSpikeLSTM.step_pipelined):
    every launch of one step (three Linear layers, two LUT preparations --
    the forget gate's for the next step, overlapped with this step's
    output-gate bootstrap --, two key switch + many-LUT bootstraps, two
    packings, the H rotation), inputs resident in HBM.  Returns a dict for
    the JSON line."""
    import numpy as np
    import torch
    sys.path.insert(0, PKG)
    import parameters_crypto as PC
    import setup_crypto as SC
    from nn_tfhe.Spike_LSTM import CipherSpikeLSTM
    from cuTFHE.multiply import check_pair_errors
    from nn_tfhe.engine import SequenceEngine
    from schemas.RLWE_jax import encrypt

    cfg = LSTM_CFG
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(cfg["seed"])
    t0 = time.perf_counter()
    sk, ksk_packing, galois_key, bsk, ksk_bs, _ = SC.sample_crypto_keys(cfg["seed"])
    keygen_s = time.perf_counter() - t0
    hid, nl, K = cfg["hidden"], cfg["n_lut"], cfg["input_dim"] + cfg["hidden"]
    G = hid // nl
    heavy, _ = SC.get_tfhe_activation(bsk, ksk_bs, ksk_packing, galois_key, np.ones(G), nl,
                                      device=dev)
    lstm, plain = synthetic_layer(rng, heavy, cfg["input_dim"], hid, nl, 128)
    out = {"workload": f"one encrypted spiking-LSTM layer: input {cfg['input_dim']}, hidden {hid}, "
                       f"n_lut {nl}, 3 gates (f, i, o: the reference's spiking cell), at the "
                       f"reference's LSTM parameters q=2^64, N=2048, n=702, l=2, Bg=2^13, c=3 "
                       f"(not the torus32 PBS parameters)",
           "keygen_s": round(keygen_s, 2), "bootstraps_per_step_per_sequence": 3 * G}
    for S in (1, 4):
        out[f"decrypt_check_batch{S}"] = lstm_decrypt_check(lstm, plain, sk, S, 3, rng)
    N = PC.degree
    for S, steps in ((1, args.lstm_steps), (args.lstm_batch, max(4, args.lstm_steps // 4))):
        if steps <= 0 or S <= 0:
            continue
        msg = np.zeros((steps + 2, N), np.int64)
        msg[:, :cfg["input_dim"]] = rng.integers(0, 2, (steps + 2, cfg["input_dim"])) << 56
        A, b = encrypt(msg, sk, PC.dict_params, rng)
        x = torch.stack([torch.as_tensor(A), torch.as_tensor(b)], dim=1).to(dev)   # (T, 2, N)
        x = x.unsqueeze(1).expand(-1, S, -1, -1).contiguous()                       # (T, S, 2, N)
        H = torch.zeros(S, 2, N, dtype=torch.int64, device=dev)
        C = torch.zeros_like(H)
        # steps chained as a sequence runs them: each also prepares the next
        # step's forget-gate LUT (overlapped with its output-gate bootstrap)
        lut = lstm.forget_lut(C)
        for t in range(2):                                                          # warm-up
            H, C, lut = lstm.step_pipelined(H, C, x[t], lut)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        for t in range(steps):
            evs[t][0].record(stream)
            H, C, lut = lstm.step_pipelined(H, C, x[t + 2], lut)
            evs[t][1].record(stream)
        torch.cuda.synchronize()
        check_pair_errors()        # a run-out pair wait would have left wrong states
        ms = np.array([a.elapsed_time(e) for a, e in evs])
        key = "latency_batch1" if S == 1 else f"throughput_batch{S}"
        out[key] = {"sequences": S, "steps": steps, "p50_ms": round(float(np.percentile(ms, 50)), 3),
                    "p99_ms": round(float(np.percentile(ms, 99)), 3),
                    "mean_ms": round(float(ms.mean()), 3),
                    "bootstraps_per_s": round(3 * G * S / (ms.mean() / 1e3), 1)}
        # the same steps replayed from a captured HIP graph (the timestep
        # engine's path): no host launch gaps between the step's kernels
        eng = SequenceEngine(lstm, use_graph=True)
        seqs = [(x[:, i, 0], x[:, i, 1]) for i in range(S)]
        st: list = []
        eng.run(seqs, step_times=st)
        gms = np.array([m for _, m in st[2:]])              # step 0 captures the graph
        out[key + "_graph"] = {"sequences": S, "steps": len(gms),
                               "p50_ms": round(float(np.percentile(gms, 50)), 3),
                               "p99_ms": round(float(np.percentile(gms, 99)), 3),
                               "mean_ms": round(float(gms.mean()), 3),
                               "bootstraps_per_s": round(3 * G * S / (gms.mean() / 1e3), 1)}
    from cuTFHE.multiply import tuned_forms
    out["tuned_forms"] = {str(k): v for k, v in tuned_forms().items()}
    return out


# full classifier: 32 sequences of lengths U[8, 64], Zipf(1.1) token ids over
# a 30522-word vocabulary, embedding 128 binarized to spikes, hidden 256,
# n_lut 4, head Linear(256 -> 32) -> ReLU bootstrap -> packing -> Linear(32 -> 2)
CLF_CFG = dict(sequences=32, len_lo=8, len_hi=64, vocab=30522, zipf=1.1, input_dim=128, hidden=256,
               n_lut=4, head_hidden=32, classes=2, seed=1)


def run_classifier(args):
    """All timesteps of 32 uneven-length sequences through the timestep engine
    (live sequences batched per step, steps replayed from captured HIP
    graphs), then the encrypted head; the timed region starts with every
    input ciphertext resident in HBM."""
    import numpy as np
    import torch
    sys.path.insert(0, PKG)
    import parameters_crypto as PC
    import setup_crypto as SC
    from nn_tfhe.Spike_LSTM import CipherMLP, CipherSpikeLSTM
    from nn_tfhe.engine import SequenceEngine
    from cuTFHE.multiply import tuned_forms
    from schemas.RLWE_jax import encrypt
    from schemas.text_jax import rotate_ciphertext

    cfg = CLF_CFG
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(cfg["seed"])
    sk, ksk_packing, galois_key, bsk, ksk_bs, _ = SC.sample_crypto_keys(cfg["seed"])
    hid, nl, K = cfg["hidden"], cfg["n_lut"], cfg["input_dim"] + cfg["hidden"]
    G = hid // nl
    heavy, relu = SC.get_tfhe_activation(bsk, ksk_bs, ksk_packing, galois_key,
                                         np.full(cfg["head_hidden"], 4.0), nl, device=dev)
    lstm, plain = synthetic_layer(rng, heavy, cfg["input_dim"], hid, nl, cfg["len_hi"])
    # int8 head weights in fixed point (x 1/64, x 1/128): the ReLU inputs stay
    # inside the LUT range [-4, 4) and the logits inside the torus
    W1 = rng.integers(-127, 128, (hid, cfg["head_hidden"]))
    W2 = rng.integers(-127, 128, (cfg["head_hidden"], cfg["classes"]))
    head1 = SC.make_linear(W1, np.zeros(cfg["head_hidden"]), 1 / 64, PC.dict_params, device=dev)
    head2 = SC.make_linear(W2, np.zeros(cfg["classes"]), 1 / 128, PC.dict_params, device=dev)
    head = CipherMLP(PC.dict_params, PC.dict_params_lut, hid, head1, head2, relu)
    Wt1, Wt2 = np.rint(W1 / 64).astype(np.int64), np.rint(W2 / 128).astype(np.int64)
    # client side: token ids, binarized embeddings, encryption
    emb = rng.standard_normal((cfg["vocab"], cfg["input_dim"])) > 0
    lens = rng.integers(cfg["len_lo"], cfg["len_hi"] + 1, cfg["sequences"])
    seqs, xs_plain = [], []
    for L_ in lens:
        tok = np.minimum(rng.zipf(cfg["zipf"], L_), cfg["vocab"]) - 1
        msg = np.zeros((L_, PC.degree), np.int64)
        msg[:, :cfg["input_dim"]] = emb[tok].astype(np.int64) << 56
        xs_plain.append(emb[tok].astype(np.int64))
        A, b = encrypt(msg, sk, PC.dict_params, rng)
        seqs.append((torch.as_tensor(A).to(dev), torch.as_tensor(b).to(dev)))
    shift = torch.full((1,), -cfg["input_dim"], dtype=torch.int64, device=dev)
    out = {"workload": f"encrypted spiking-LSTM classifier: {cfg['sequences']} sequences, lengths "
                       f"U[{cfg['len_lo']},{cfg['len_hi']}], Zipf({cfg['zipf']}) tokens over "
                       f"{cfg['vocab']}, embedding {cfg['input_dim']} -> spikes, hidden {hid}, "
                       f"n_lut {nl}, 3 gates, head {hid}->{cfg['head_hidden']}->ReLU->"
                       f"{cfg['classes']} (int8 weights x 1/64, x 1/128), q=2^64, N=2048",
           "total_timesteps": int(lens.sum()),
           "bootstraps": int(lens.sum()) * 3 * G + cfg["sequences"] * cfg["head_hidden"]}
    for use_graph in (False, True):
        eng = SequenceEngine(lstm, use_graph=use_graph)
        eng.run(seqs)                                     # warm-up: tuning, graph capture
        torch.cuda.synchronize()
        steps: list = []
        t0 = time.perf_counter()
        H, C = eng.run(seqs, step_times=steps)
        Hh = rotate_ciphertext((H[:, 0], H[:, 1]), shift)
        logits = head(Hh)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = np.array([m for _, m in steps])
        out["graph" if use_graph else "eager"] = {
            "wall_s": round(wall, 3), "bootstraps_per_s": round(out["bootstraps"] / wall, 1),
            "step_p50_ms": round(float(np.percentile(ms, 50)), 3),
            "step_p99_ms": round(float(np.percentile(ms, 99)), 3), "steps": len(ms)}
    out["logits_shape"] = list(logits[0].shape)
    out["decrypt_check"] = classifier_decrypt_check(lstm, plain, head, Wt1, Wt2, sk, seqs, lens,
                                                    xs_plain, H, logits, cfg)
    out["tuned_forms"] = {str(k): v for k, v in tuned_forms().items()}
    return out


def classifier_decrypt_check(lstm, plain, head, Wt1, Wt2, sk, seqs, lens, xs_plain, H, logits, cfg):
    """(a) The two shortest sequences re-run step by step (eager), every
    step's states decrypted against the plaintext rule, and their final
    states equal to the engine's; (b) every sequence's logits decrypted and
    compared with the plaintext head applied to its decrypted final hidden
    state (ReLU ramp within the noise bound, logits within the bound
    propagated through the second Linear, predicted class)."""
    import numpy as np
    import torch
    import parameters_crypto as PC
    from nn_tfhe.plaintext import fold, lut_index_noise, relu_ramp
    from schemas.LWE_jax import lwe_phase
    from schemas.text_jax import rotate_ciphertext
    inp, hid = cfg["input_dim"], cfg["hidden"]
    margin, _ = lstm_margin(cfg["n_lut"])
    det = ok = tot = 0
    same = True
    for i in np.argsort(lens)[:2]:
        (H1, C1), (Hs, Cs) = lstm(seqs[i])
        same &= bool(torch.equal(H[i, 0], H1[0]) and torch.equal(H[i, 1], H1[1]))
        Hd = decrypt_values((Hs[0].cpu().numpy(), Hs[1].cpu().numpy()), sk, inp, inp + hid)
        Cd = decrypt_values((Cs[0].cpu().numpy(), Cs[1].cpu().numpy()), sk, 0, hid)
        d, o, t = count_determined(plain, margin, xs_plain[i], Hd, Cd)
        det, ok, tot = det + d, ok + o, tot + t
    # head: plaintext on the decrypted final H
    shift = torch.full((1,), -inp, dtype=torch.int64, device=H.device)
    Hr = rotate_ciphertext((H[:, 0], H[:, 1]), shift)
    Hv = decrypt_values((Hr[0].cpu().numpy(), Hr[1].cpu().numpy()), sk, 0, hid)
    pre = (Hv @ Wt1) / LSTM_W_SCALE
    step = LSTM_S / (PC.degree_lut // 2)
    sig = lut_index_noise(PC.dict_params_lut, PC.dict_params_ks_LWE, PC.degree, PC.q, 1) * step
    ok_relu = np.abs(np.abs(fold(pre, LSTM_S)) - LSTM_S) >= MARGIN_SIGMAS * sig
    y_plain = relu_ramp(pre, LSTM_S, PC.degree_lut // 2)
    lg_plain = y_plain @ Wt2
    lg = lwe_phase((logits[0].cpu().numpy(), logits[1].cpu().numpy()), sk, PC.dict_params) / PC.beta_x
    lg_sig = sig * np.sqrt((Wt2 ** 2).sum(0))
    rows = ok_relu.all(axis=1)
    close = np.abs(lg - lg_plain) <= MARGIN_SIGMAS * lg_sig
    return {"sequences_stepped": 2, "units": tot, "determined_frac": round(det / max(tot, 1), 4),
            "determined_correct_frac": round(ok / max(det, 1), 6),
            "engine_equals_eager": same,
            "logits_rows_determined": int(rows.sum()),
            "logits_within_noise_frac": round(float(close[rows].mean()) if rows.any() else 0.0, 4),
            "class_agreement_frac": round(float((lg.argmax(1) == lg_plain.argmax(1))[rows].mean())
                                          if rows.any() else 0.0, 4),
            "relu_noise_sigma": round(float(sig), 4)}


# serving: Poisson arrivals of single-sequence requests (lengths U[4, 128])
# from 64 clients with their own keys; one LSTM layer (input 128, hidden 128)
SERVE_CFG = dict(clients=64, reserve_gib=24, qps=4.0, duration_s=30.0, len_lo=4, len_hi=128,
                 input_dim=128, hidden=128, n_lut=4, max_active=96, seed=2)


def run_serving(args):
    """Continuous-batching server at a fixed request rate; every client's keys
    generated on the device up front (client registration), a key pool of
    `slots` clients resident at once (LRU eviction beyond that)."""
    import numpy as np
    import torch
    sys.path.insert(0, PKG)
    import parameters_crypto as PC
    import setup_crypto as SC
    from nn_tfhe.serving import EncryptedLSTMServer, KeyPool, Request
    from schemas.RLWE_jax import encrypt

    cfg = SERVE_CFG
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(cfg["seed"])
    t0 = time.perf_counter()
    keys, secrets = {}, {}
    for c in range(cfg["clients"]):
        keys[c], secrets[c] = SC.sample_client_keys_device(1000 + c, device=dev)
    torch.cuda.synchronize()
    keygen_s = time.perf_counter() - t0
    hid, nl, K = cfg["hidden"], cfg["n_lut"], cfg["input_dim"] + cfg["hidden"]
    G = hid // nl
    factor = PC.q / 16 / 16 / PC.beta_x
    lin = [SC.make_linear(rng.integers(-7, 8, (K, G)), np.zeros(G), factor, PC.dict_params,
                          device=dev) for _ in range(3)]
    s4 = np.full(G, 4.0)
    gam = [rng.uniform(-2, 2, (nl, G)) for _ in range(3)]
    lut_f, lut_i, lut_o = SC.get_lut(nl, hid, PC.degree_lut, s4, s4, s4, gam[0], gam[1], gam[2],
                                     np.ones((nl, G)))
    margs = dict(dict_params=PC.dict_params, dict_params_lut=PC.dict_params_lut,
                 dict_params_ks_LWE=PC.dict_params_ks_LWE, dict_params_packing=PC.dict_params_packing,
                 n_lut=nl, input_dim=cfg["input_dim"], hidden=hid, max_len=cfg["len_hi"],
                 linear_f=lin[0], linear_i=lin[1], linear_o=lin[2], lut_f=lut_f, lut_i=lut_i,
                 lut_o=lut_o, beta_x=PC.beta_x)
    # admission by HBM footprint: as many resident key sets as free device
    # memory holds beside a reserve for the step workspaces, at most one per client
    slot_bytes = sum(v.numel() * v.element_size() for v in keys[0].values())
    free_b, _ = torch.cuda.mem_get_info(dev)
    slots = int(max(1, min(cfg["clients"], (free_b - cfg["reserve_gib"] * 2 ** 30) // slot_bytes)))
    pool = KeyPool(slots, keys[0], device=dev)
    srv = EncryptedLSTMServer(margs, pool, lambda c: keys[c], max_active=cfg["max_active"])
    # warm-up: one short request per batch-size bucket tunes the launch shapes
    for S in (1, 2, 4, 8, 16, 32, 64, 96):
        srv.lstm.heavyBS.set_slots(torch.zeros(S, dtype=torch.int32))
        z = torch.zeros(S, 2, PC.degree, dtype=torch.int64, device=dev)
        srv.lstm.step(z, z, z)
    torch.cuda.synchronize()
    # client side: Poisson arrivals, encrypted spike frames
    reqs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / cfg["qps"])
        if t > cfg["duration_s"]:
            break
        c = int(rng.integers(0, cfg["clients"]))
        L_ = int(rng.integers(cfg["len_lo"], cfg["len_hi"] + 1))
        msg = np.zeros((L_, PC.degree), np.int64)
        msg[:, :cfg["input_dim"]] = rng.integers(0, 2, (L_, cfg["input_dim"])) << 56
        A, b = encrypt(msg, secrets[c][0], PC.dict_params, rng)
        reqs.append(Request(len(reqs), c, torch.stack([torch.as_tensor(A), torch.as_tensor(b)],
                                                       dim=1).to(dev), t))
    w0 = time.perf_counter()
    done = srv.run(reqs)
    wall = time.perf_counter() - w0
    check = serving_decrypt_check(done, secrets, margs, keys, pool, dev)
    lat = np.array([r.done - r.arrival for r in done])
    per_step = np.array([(r.done - r.arrival) / r.x.shape[0] for r in done])
    return {"workload": f"continuous-batching server: Poisson {cfg['qps']} req/s for "
                        f"{cfg['duration_s']} s, {cfg['clients']} clients with own keys, "
                        f"{slots} key slots (free HBM / {slot_bytes / 2 ** 20:.0f} MiB per "
                        f"client), lengths U[{cfg['len_lo']},{cfg['len_hi']}], "
                        f"one LSTM layer input {cfg['input_dim']} hidden {hid}, q=2^64, N=2048",
            "requests": len(done), "timesteps": int(sum(r.x.shape[0] for r in done)),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
            "p50_latency_per_timestep_ms": round(float(np.percentile(per_step, 50)) * 1e3, 2),
            "p99_latency_per_timestep_ms": round(float(np.percentile(per_step, 99)) * 1e3, 2),
            "bootstraps_per_s": round(srv.bootstraps / wall, 1), "wall_s": round(wall, 2),
            "ticks": srv.ticks, "gpu_busy_frac": round(srv.gpu_busy_s() / wall, 3),
            "key_loads": pool.loads, "decrypt_check": check,
            "key_bytes_per_client": pool.slot_bytes, "key_pool_bytes": pool.footprint_bytes(),
            "client_keygen_s": round(keygen_s, 2)}


def serving_decrypt_check(done, secrets, margs, keys, pool, dev):
    """Every finished request's final cell state decrypted under its own
    client's key: integer values (units of beta_x) of magnitude at most the
    sequence length (the cell state counts input spikes); and the three
    shortest requests served again alone, final states compared bit for bit
    with the mixed run."""
    import numpy as np
    import torch
    import parameters_crypto as PC
    from nn_tfhe.serving import EncryptedLSTMServer, KeyPool, Request
    from schemas.RLWE_jax import _phase
    hid = margs["hidden"]
    ok = 0
    for r in done:
        ph = _phase((r.result[1][0].cpu().numpy(), r.result[1][1].cpu().numpy()),
                    secrets[r.client][0], PC.dict_params)[:hid] / PC.beta_x
        ok += int(np.abs(ph - np.rint(ph)).max() < 0.05 and
                  np.abs(np.rint(ph)).max() <= r.x.shape[0])
    same = 0
    short = sorted(done, key=lambda r: r.x.shape[0])[:3]
    for r in short:
        solo_pool = KeyPool(1, keys[r.client], device=dev)
        solo = EncryptedLSTMServer(margs, solo_pool, lambda c: keys[c], max_active=1)
        fin = solo.run([Request(r.rid, r.client, r.x, 0.0)], clock=lambda: float(solo.ticks))
        same += int(torch.equal(fin[0].result[0], r.result[0]) and
                    torch.equal(fin[0].result[1], r.result[1]))
    return {"requests": len(done), "cell_states_integral_frac": round(ok / max(len(done), 1), 4),
            "alone_equals_mix": f"{same}/{len(short)}"}


def cpu_baseline_child(sample: int, seed: int = 1):
    """Oracle (NumPy FP64 restatement of the reference algorithm) on a
    bounded sample of the same workload, single-threaded."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import tfhe_oracle as O
    cfg = CFG
    rng = np.random.default_rng(seed)
    N, n, L, c = cfg["N"], cfg["n"], cfg["L"], cfg["collapse"]
    n_iter = n // c
    m = 2 ** c
    # random FFT-domain key of the right shape: the timing does not depend on
    # the key values
    bsk_k = rng.standard_normal((n_iter, 2, L, 2, m, N // 2, 2)).view(np.complex128)[..., 0]
    bsk_k *= 2.0 ** 30
    a = rng.integers(-2 ** 31, 2 ** 31, size=(sample, n))
    b = rng.integers(-2 ** 31, 2 ** 31, size=sample)
    lut = rng.integers(-2 ** 31, 2 ** 31, size=(sample, 2, N))
    t0 = time.perf_counter()
    idx = O.rotation_indices(a, 32, N, c, cfg["n_lut"])
    br = O.blind_rotate_f64(lut, bsk_k, idx, cfg["log_beta"], 32, L)
    O.cuboot_merge_tail(br, b, 32, cfg["n_lut"])
    dt = time.perf_counter() - t0
    print(json.dumps({"value": sample / dt, "seconds": dt}))


def cpu_config1_child(sample: int, seed: int = 1):
    """One worker of the config-1 all-cores figure: the oracle's NumPy
    restatement of the PBS, extraction and key switch of `sample`
    ciphertexts at n=630, N=1024, single-threaded, on random keys and inputs
    of the same shapes (the time does not depend on their values; the
    one-core run checks the decryptions); two warm-ups, then the median of
    five repetitions."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import tfhe_oracle as O
    c = PBS_KS_CFG
    n, N, L, lb, cl, lq = c["n"], c["N"], c["L"], c["log_beta"], c["collapse"], c["log_q"]
    rng = np.random.default_rng(seed)
    fft_k = rng.standard_normal((n // cl, 2, L, 2, 2 ** cl, N // 2, 2)).view(np.complex128)[..., 0]
    fft_k *= 2.0 ** 30
    ka = rng.integers(-2 ** 31, 2 ** 31, size=(c["ks_l"], N, n))
    kb = rng.integers(-2 ** 31, 2 ** 31, size=(c["ks_l"], N))
    a = rng.integers(-2 ** 31, 2 ** 31, size=(sample, n))
    b = rng.integers(-2 ** 31, 2 ** 31, size=sample)
    lut = rng.integers(-2 ** 31, 2 ** 31, size=(2, N))

    def pbs(lo, hi):
        idx = O.rotation_indices(a[lo:hi], lq, N, cl, 1)
        br = O.blind_rotate_f64(np.broadcast_to(lut, (hi - lo, 2, N)), fft_k, idx, lb, lq, L)
        m_, b_ = O.cuboot_merge_tail(br, b[lo:hi], lq, 1)
        O.key_switch_lwe(ka, kb, m_[:, 0], b_[:, 0], lq, c["ks_log_beta"], c["ks_l"])

    for _ in range(2):                          # warm-ups
        pbs(0, min(4, sample))
    reps = []
    for _ in range(5):
        t0 = time.perf_counter()
        pbs(0, sample)
        reps.append(time.perf_counter() - t0)
    dt = float(np.median(reps))
    print(json.dumps({"value": sample / dt, "seconds": dt}))


# the reference's CPU-runnable case: 256 LWE ciphertexts, n=630, N=1024, l=3,
# log2 Bg=7, collapse 3, identity LUT on 3-bit messages with a padding bit,
# then sample extract and key switch back to n (base 2^4, l=5)
PBS_KS_CFG = dict(B=256, n=630, N=1024, L=3, log_beta=7, collapse=3, msg_bits=3, log_q=32,
            sigma_lwe=2.0 ** 17, sigma_bsk=2.0 ** 7, ks_log_beta=4, ks_l=5, sigma_ks=2.0 ** 10,
            cpu_sample=256)


def run_pbs_keyswitch(cpu: bool):
    """PBS (blind rotation, extraction, key switch) of 256 ciphertexts at N=1024
    on the GPU through the product path, decrypted and compared with the
    messages; with cpu=True also the oracle's NumPy restatement of the same
    computation, timed on one host core (the reference's JAX-CPU path cannot
    run here).  The oracle also builds the keys and inputs (seeded)."""
    import numpy as np
    import torch
    sys.path.insert(0, ROOT)
    sys.path.insert(0, PKG)
    from cuTFHE.multiply import mul_bootstrap, rotation_indices
    from oracle import tfhe_oracle as O
    from schemas.LWE_jax import key_switch_LWE
    from schemas.RLWE_jax import rotate_extract

    c = PBS_KS_CFG
    B, n, N, L, lb, cl, lq = c["B"], c["n"], c["N"], c["L"], c["log_beta"], c["collapse"], c["log_q"]
    rng = np.random.default_rng(0)
    s = rng.integers(0, 2, n)
    S = rng.integers(0, 2, N)
    bsk = O.bsk_collapse_coef(rng, s, S, lq, lb, L, cl, c["sigma_bsk"])
    fft_k = np.ascontiguousarray(O.reshape_bsk_for_kernel(O.bsk_to_fourier_ordered(bsk), n // cl, cl))
    del bsk
    ka, kb = O.keyswitch_key(rng, s, S, lq, c["ks_log_beta"], c["ks_l"], c["sigma_ks"])
    delta = 1 << (lq - c["msg_bits"] - 1)
    msg = rng.integers(0, 1 << c["msg_bits"], B)
    a, b = O.lwe_encrypt(rng, msg * delta, s, lq, c["sigma_lwe"])
    W = N >> c["msg_bits"]
    j = np.arange(N)
    T = delta * (((j + W // 2) // W) % (1 << c["msg_bits"]))
    lut = np.stack([np.zeros(N, np.int64), np.asarray(T, np.int64)])

    dev = torch.device("cuda:0")
    pk = torch.as_tensor(a.astype(np.float64)).to(dev)
    bd = torch.as_tensor(b.astype(np.float64)).reshape(B, 1).to(dev)
    bsk_t = torch.as_tensor(fft_k).to(dev)
    lut_t = torch.as_tensor(np.ascontiguousarray(np.broadcast_to(lut, (B, 2, N)))).to(dev)
    ksk = (torch.as_tensor(ka.astype(np.float64)).to(dev),
           torch.as_tensor(kb.astype(np.float64))[..., None].to(dev))
    kp = {"beta_ks": 2 ** c["ks_log_beta"], "l_ks": c["ks_l"], "q": 2.0 ** lq}

    def pbs():
        idx = rotation_indices(pk, lq, N, cl, 1)
        br = mul_bootstrap(lut_t, bsk_t, None, idx, lb, lq, L, form="auto")
        m_, b_ = rotate_extract(br, bd, lq, 1)
        return key_switch_LWE(ksk, (m_[:, 0], b_[:, 0]), kp)

    pbs()
    torch.cuda.synchronize()
    times = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ga, gb = pbs()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    ga = ga.cpu().numpy().astype(np.int64)
    gb = gb.cpu().numpy().astype(np.int64).reshape(-1)
    ph = O.centered_mod(gb - ga @ s.astype(np.int64), lq)
    dec = np.mod(np.rint(ph / delta).astype(np.int64), 1 << (c["msg_bits"] + 1))
    ms = float(np.median(times))
    out = {"workload": f"PBS of {B} LWE ciphertexts, n={n}, N={N}, l={L}, log2 Bg={lb}, "
                       f"collapse {cl}, identity LUT, 3-bit messages, then sample extract and key "
                       f"switch {N}->{n} (base 2^{c['ks_log_beta']}, l={c['ks_l']})",
           "gpu": {"ms": round(ms, 3), "bootstraps_per_s": round(B / (ms / 1e3), 1),
                   "messages_correct": float((dec == msg).mean())}}
    if cpu:
        from threadpoolctl import threadpool_limits
        Bc = c["cpu_sample"]                    # a bounded sample: the first Bc ciphertexts
        rep = Bc // 4                           # one repetition: a quarter of the sample

        def restate(lo, hi):
            idx = O.rotation_indices(a[lo:hi], lq, N, cl, 1)
            br = O.blind_rotate_f64(np.broadcast_to(lut, (hi - lo, 2, N)), fft_k, idx, lb, lq, L)
            m_, b_ = O.cuboot_merge_tail(br, b[lo:hi], lq, 1)
            return O.key_switch_lwe(ka, kb, m_[:, 0], b_[:, 0], lq, c["ks_log_beta"], c["ks_l"])

        with threadpool_limits(limits=1):
            for _ in range(2):                  # warm-ups
                restate(0, 4)
            reps, parts = [], []
            for k in range(5):                  # the four quarters, then the first again
                lo = (k % 4) * rep
                t0 = time.perf_counter()
                r_ = restate(lo, lo + rep)
                reps.append(time.perf_counter() - t0)
                if k < 4:
                    parts.append(r_)
        dt = float(np.median(reps))
        ra = np.concatenate([np.asarray(p_[0], np.int64) for p_ in parts])
        rb = np.concatenate([np.asarray(p_[1], np.int64).reshape(-1) for p_ in parts])
        ph = O.centered_mod(rb - ra @ s, lq)
        dec_c = np.mod(np.rint(ph / delta).astype(np.int64), 1 << (c["msg_bits"] + 1))
        out["cpu_restatement"] = {"sample": Bc, "timing": f"median of 5 repetitions of {rep} "
                                  f"ciphertexts after 2 warm-ups",
                                  "seconds_per_repetition": round(dt, 3),
                                  "bootstraps_per_s": round(rep / dt, 2), "cores": 1,
                                  "messages_correct": float((dec_c == msg[:Bc]).mean()),
                                  "gpu_equals_restatement": bool(np.array_equal(dec[:Bc], dec_c)),
                                  "cpu_model": cpu_model()}
        # all cores: one single-threaded worker per host core at once
        P, per = host_cores(), 16
        try:
            many = run_cpu_workers("--cpu-config1-child", P, per)
            wall = max(d["seconds"] for d in many)
            out["cpu_restatement"]["all_cores"] = {
                "bootstraps_per_s": round(P * per / wall, 2), "cores": P,
                "sample": f"{per} ciphertexts per process, {P} processes at once, each the "
                          f"median of 5 repetitions after 2 warm-ups; slowest {wall:.2f} s "
                          f"(random keys and inputs of the same shapes)"}
        except (RuntimeError, subprocess.TimeoutExpired) as e:
            out["cpu_restatement"]["all_cores"] = {"error": str(e)}
    return out


def precision_check(sample: int, restate_sample: int, oracle_sample: int):
    """Phase error of the FP64 blind rotation (batched form) against exact
    products, on `sample` ciphertexts of the batched-PBS workload with real
    seeded keys.  The exact products at this size come from the split form
    (exact integer arithmetic); the int128 oracle re-checks the first
    `oracle_sample` of them, and the NumPy FP64 restatement of the reference
    algorithm (the reference's own arithmetic, for its error) runs on the
    first `restate_sample` (all host cores).  The oracle and the restatement
    are the checkers; the GPU kernels are what is measured."""
    import numpy as np
    import torch
    sys.path.insert(0, ROOT)
    sys.path.insert(0, PKG)
    from cuTFHE.multiply import mul_bootstrap, rotation_indices
    from oracle import exact as X
    from oracle import tfhe_oracle as O
    from schemas.RLWE_jax import rotate_extract

    cfg = CFG
    N, n, L, lb, c = cfg["N"], cfg["n"], cfg["L"], cfg["log_beta"], cfg["collapse"]
    nl, mb, lq = cfg["n_lut"], cfg["msg_bits"], cfg["log_q"]
    n_iter = n // c
    t0 = time.perf_counter()
    rng = np.random.default_rng(5)
    s = rng.integers(0, 2, n)
    S = rng.integers(0, 2, N)
    bsk = O.bsk_collapse_coef(rng, s, S, lq, lb, L, c, cfg["sigma_bsk"])
    fft_k = np.ascontiguousarray(O.reshape_bsk_for_kernel(O.bsk_to_fourier_ordered(bsk), n_iter, c))
    coef_k = np.ascontiguousarray(O.bsk_coef_kernel_layout(bsk, n_iter, c))
    del bsk
    delta = 1 << (lq - mb - 1)
    msg = rng.integers(-(1 << (mb - 1)), 1 << (mb - 1), sample)
    T = O.lut_from_function(N, nl, MANY_LUT_FNS[:nl], mb, delta)
    a, b = O.lwe_encrypt(rng, msg * delta, s, lq, cfg["sigma_lwe"])
    lut = np.stack([np.zeros(N, np.int64), np.asarray(T, np.int64)])
    lut_b = np.ascontiguousarray(np.broadcast_to(lut, (sample, 2, N)))

    dev = torch.device("cuda:0")
    pk = torch.as_tensor(a.astype(np.float64)).to(dev)
    bd = torch.as_tensor(b.astype(np.float64)).reshape(sample, 1).to(dev)
    bsk_t = torch.as_tensor(fft_k).to(dev)
    lut_t = torch.as_tensor(lut_b).to(dev)
    idx_t = rotation_indices(pk, lq, N, c, nl)
    gpu = {}
    for form in ("batched", "plain", "split"):
        br = mul_bootstrap(lut_t, bsk_t, None, idx_t, lb, lq, L, form=form)
        masks, bodies = rotate_extract(br, bd, lq, nl)
        torch.cuda.synchronize()
        gpu[form] = (br.cpu().numpy(), masks.cpu().numpy().astype(np.int64),
                     bodies.cpu().numpy().astype(np.int64).reshape(sample, nl))
    del bsk_t
    idx = O.rotation_indices(a, lq, N, c, nl)
    br_x = X.blind_rotate(lut, coef_k, idx[:oracle_sample], lb, lq, L)
    br_f = restate_blind_rotate(lut, fft_k, idx[:restate_sample], lb, lq, L)

    def extract(br, k):
        m_, b_ = O.cuboot_merge_tail(br, b[:k], lq, nl)
        return np.asarray(m_, np.int64), np.asarray(b_, np.int64).reshape(k, nl)

    expect = np.array([[f(int(x)) for f in MANY_LUT_FNS[:nl]] for x in msg], np.int64)
    mod = 1 << (mb + 1)

    def decrypt(masks, bodies):
        k = masks.shape[0]
        ph = O.centered_mod(bodies - masks @ S.astype(np.int64), lq)
        return (np.mod(np.rint(ph / delta).astype(np.int64), mod),
                O.centered_mod(ph - expect[:k] * delta, lq))

    dec_g, err_g = decrypt(*gpu["batched"][1:])
    dec_s, err_s = decrypt(*gpu["split"][1:])
    dec_f, err_f = decrypt(*extract(br_f, restate_sample))
    dev_g = O.centered_mod(gpu["batched"][0] - gpu["split"][0], lq)
    dev_f = O.centered_mod(br_f - gpu["split"][0][:restate_sample], lq)
    pred = O.predicted_bs_std(N, n, lb, L, c, cfg["sigma_bsk"], lq)
    R = restate_sample
    return {
        "sample": f"{sample} ciphertexts of the batched-PBS workload, seeded keys "
                  f"(sigma_bsk 2^{math.log2(cfg['sigma_bsk']):.0f}, sigma_lwe "
                  f"2^{math.log2(cfg['sigma_lwe']):.0f}); exact products from the split form, "
                  f"its first {oracle_sample} re-checked by the int128 oracle; the NumPy FP64 "
                  f"restatement of the reference algorithm on the first {R}",
        "split_equals_int128_oracle": bool(np.array_equal(gpu["split"][0][:oracle_sample], br_x)),
        "coef_dev_vs_exact": {"gpu_max_abs": int(np.abs(dev_g).max()), "gpu_std": float(dev_g.std()),
                              "restatement_max_abs": int(np.abs(dev_f).max()),
                              "restatement_std": float(dev_f.std())},
        "batched_equals_plain": bool(np.array_equal(gpu["batched"][0], gpu["plain"][0])),
        "phase_err_std": {"gpu": float(err_g.std()), "exact": float(err_s.std()),
                          "gpu_on_restatement_sample": float(err_g[:R].std()),
                          "restatement": float(err_f.std()), "predicted_bs": pred},
        "phase_err_std_ratio_gpu_vs_restatement": float(err_g[:R].std() / err_f.std()),
        "phase_err_std_ratio_gpu_vs_exact": float(err_g.std() / err_s.std()),
        "messages_identical_gpu_vs_exact": float((dec_g == dec_s).mean()),
        "messages_identical_gpu_vs_restatement": float((dec_g[:R] == dec_f).mean()),
        "messages_equal_lut": float((dec_g == np.mod(expect, mod)).mean()),
        "seconds": round(time.perf_counter() - t0, 1),
    }


def _restate_child(path: str, lo: int, hi: int, lb: int, lq: int, L: int):
    """Worker of restate_blind_rotate: rows lo:hi of the inputs saved at path."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import tfhe_oracle as O
    d = np.load(path)
    lut = np.ascontiguousarray(np.broadcast_to(d["lut"], (hi - lo,) + d["lut"].shape[-2:]))
    br = np.asarray(O.blind_rotate_f64(lut, d["bsk"], d["idx"][lo:hi], lb, lq, L), np.int64)
    np.save(path.replace(".npz", f".{lo}.npy"), br)


def restate_blind_rotate(lut, fft_k, idx, lb, lq, L):
    """The oracle's NumPy FP64 restatement of the reference's blind rotation
    over all host cores (one single-threaded process per core)."""
    import numpy as np
    import tempfile
    P = host_cores()
    B = idx.shape[0]
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "restate.npz")
        np.savez(path, lut=lut, bsk=fft_k, idx=idx)
        bounds = [(B * i // P, B * (i + 1) // P) for i in range(P)]
        procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--restate-child",
                                   f"{path},{lo},{hi},{lb},{lq},{L}"], env=single_thread_env())
                 for lo, hi in bounds if hi > lo]
        for p in procs:
            if p.wait() != 0:
                raise RuntimeError("restatement worker failed")
        return np.concatenate([np.load(path.replace(".npz", f".{lo}.npy"))
                               for lo, hi in bounds if hi > lo])


def host_cores() -> int:
    """Host cores this process may use (at most 16, the GPU box's share)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def single_thread_env() -> dict:
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    return env


def run_cpu_workers(flag: str, n_procs: int, per: int) -> list:
    """n_procs single-threaded worker processes of bench.py (`flag` = the
    worker's switch), started together, `per` ciphertexts each; returns
    their JSON lines."""
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), flag, str(per),
                               "--cpu-baseline-seed", str(1 + i)],
                              env=single_thread_env(), stdout=subprocess.PIPE, text=True)
             for i in range(n_procs)]
    outs = []
    for p in procs:
        o, _ = p.communicate(timeout=900)
        if p.returncode != 0:
            raise RuntimeError("CPU baseline worker failed")
        outs.append(json.loads(o.strip().splitlines()[-1]))
    return outs


def cpu_baseline(sample: int):
    """The oracle's NumPy FP64 restatement of the reference algorithm on a
    bounded sample of the headline workload: one single-threaded process
    (1 core), then one process per host core at once (all cores)."""
    try:
        one = run_cpu_workers("--cpu-baseline-child", 1, sample)[0]
        P = host_cores()
        per = max(8, sample // 4)
        many = run_cpu_workers("--cpu-baseline-child", P, per)
        wall = max(d["seconds"] for d in many)
        all_rate = P * per / wall
    except (RuntimeError, subprocess.TimeoutExpired) as e:
        return {"error": str(e)}
    return {"value": round(one["value"], 4), "unit": "bootstraps/s", "cores": 1, "kind": "port",
            "all_cores": {"value": round(all_rate, 3), "cores": P,
                          "sample": f"{per} ciphertexts per process, {P} processes at once, "
                                    f"slowest {wall:.1f} s"},
            "cpu_model": cpu_model(),
            "sample": f"{sample} ciphertexts of the same workload through the NumPy FP64 "
                      f"restatement of the reference algorithm (oracle), {one['seconds']:.1f} s, "
                      f"1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=CFG["B"])
    ap.add_argument("--form", choices=("plain", "batched", "both"), default="both",
                    help="blind-rotation form(s) to time; the JSON value is the batched form "
                         "unless only plain is asked for")
    ap.add_argument("--cpu-sample", type=int, default=128)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one elementwise negation over this many MiB (reads and "
                         "writes that many bytes): the calibration kernel for HBM counters")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision-sample", type=int, default=4096,
                    help="ciphertexts whose FP64 blind rotation is compared with exact products "
                         "(phase-error statistics; 0: skip)")
    ap.add_argument("--precision-restate", type=int, default=256,
                    help="of those, ciphertexts also run through the NumPy FP64 restatement of "
                         "the reference algorithm (all host cores)")
    ap.add_argument("--precision-oracle", type=int, default=16,
                    help="of those, ciphertexts re-checked by the int128 oracle")
    ap.add_argument("--lstm-steps", type=int, default=32,
                    help="timesteps of the encrypted spiking-LSTM layer to time (0: skip)")
    ap.add_argument("--lstm-batch", type=int, default=32,
                    help="sequences per timestep for the LSTM throughput line")
    ap.add_argument("--classifier", type=int, default=1,
                    help="run the 32-sequence encrypted classifier (0: skip)")
    ap.add_argument("--serving", type=int, default=1,
                    help="run the continuous-batching serving workload (0: skip)")
    ap.add_argument("--pbs-keyswitch", type=int, default=1,
                    help="run the 256-ciphertext N=1024 PBS + key switch case (0: skip)")
    ap.add_argument("--form-sweep", type=int, default=1,
                    help="time the blind-rotation launch forms at several batch sizes (0: skip)")
    ap.add_argument("--split", type=int, default=1,
                    help="also time the split form (exact int8 MFMA) at the headline workload")
    ap.add_argument("--cpu-baseline-child", type=int, default=0)
    ap.add_argument("--cpu-config1-child", type=int, default=0)
    ap.add_argument("--cpu-baseline-seed", type=int, default=1)
    ap.add_argument("--restate-child", type=str, default="")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        cpu_baseline_child(args.cpu_baseline_child, args.cpu_baseline_seed)
        return
    if args.cpu_config1_child:
        cpu_config1_child(args.cpu_config1_child, args.cpu_baseline_seed)
        return
    if args.restate_child:
        p, lo, hi, lb, lq, L = args.restate_child.split(",")
        _restate_child(p, int(lo), int(hi), int(lb), int(lq), int(L))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the engine has no "
              f"multi-GPU path", file=sys.stderr)
        sys.exit(2)
    res = run_gpu(args) if args.steps > 0 else {}
    if args.form_sweep:
        res["form_sweep"] = run_form_sweep()
    if args.lstm_steps > 0:
        res["lstm"] = run_lstm(args)
    if args.classifier:
        res["classifier"] = run_classifier(args)
    if args.serving:
        res["serving"] = run_serving(args)
    if args.copy_mib:
        import torch
        src = torch.ones(args.copy_mib * 2 ** 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        res["copy_calibration_bytes"] = args.copy_mib * 2 ** 20
    res["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(args.cpu_sample)
    if not args.no_cpu_baseline and args.precision_sample > 0:
        res["precision"] = precision_check(args.precision_sample,
                                           min(args.precision_restate, args.precision_sample),
                                           min(args.precision_oracle, args.precision_sample))
    if args.pbs_keyswitch:
        res["pbs_keyswitch_n1024"] = run_pbs_keyswitch(cpu=not args.no_cpu_baseline)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
