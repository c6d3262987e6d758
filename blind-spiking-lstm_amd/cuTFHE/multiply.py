"""Drop-in for the reference's cuTFHE/multiply.py on a PyTorch-ROCm host.

Same function names, argument order and meaning as cuTFHE/multiply.py:17-92;
arrays are torch tensors on the GPU instead of jax arrays.  Every call runs
hand-written HIP kernels through the C ABI (include/bslstm_tfhe.h) on the
tensor's current stream; nothing falls back to the CPU.
"""
from __future__ import annotations

import math

import torch

from . import _core


def _as_f64(x: torch.Tensor) -> torch.Tensor:
    return x if x.dtype == torch.float64 else x.to(torch.float64)


def rotation_indices(public_key: torch.Tensor, log_q: int, degree: int, collapse: int,
                     n_lut: int = 1) -> torch.Tensor:
    """Rotation exponents of cuTFHE/multiply.py:67-79 on the GPU.

    public_key (..., n) -> int32 (..., n//collapse, 2**collapse):
    idx = int32(rint(2N*val/(q*n_lut))*n_lut) mod 2N, val = pk groups . bits.
    """
    _core.require_device(public_key)
    pk = _as_f64(public_key).contiguous()
    n = pk.shape[-1]
    if n % collapse:
        raise _core.TFHEInvalidArgument(f"n={n} is not a multiple of collapse={collapse}")
    batch = pk.numel() // n if n else 0
    out = torch.empty(*pk.shape[:-1], n // collapse, 2 ** collapse, dtype=torch.int32,
                      device=pk.device)
    _core.check(_core.lib().bslstm_rotation_indices(_core.stream_of(pk), pk.data_ptr(), batch, n,
                                                    collapse, log_q, degree, n_lut,
                                                    out.data_ptr()))
    return out


def prepare_bsk(BSK: torch.Tensor, n: int, collapse: int) -> torch.Tensor:
    """BSK reshape/permute of cuTFHE/multiply.py:81-86, materialised once.

    (2[pk/b], 2[L/R], l, n_eq, N/2) complex128 -> contiguous
    (n_iter, 2[a], l, 2[L/R], 2^collapse, N/2).  A 6-D tensor is taken to be
    in kernel layout already and is returned (contiguous) as is.
    """
    if BSK.dim() == 6:
        return BSK.contiguous()
    n_iter = n // collapse
    shape = list(BSK.shape)
    shape[3] = n_iter
    shape.insert(4, 2 ** collapse)
    return BSK.reshape(shape).permute(3, 0, 2, 1, 4, 5).contiguous()


FORMS = ("plain", "batched", "rows", "pair")

# polls of a pair-form workgroup waiting on its partner before it gives up
# (each poll sleeps ~50 ns and reads the flag, ~0.5-2 s in all; a launch whose
# workgroups are all resident never comes near this)
PAIR_SPIN_LIMIT = 1 << 20
PAIR_RUN_OUT = ("a pair-form blind rotation gave up waiting for its partner workgroup "
                "(the device did not keep both workgroups of a pair resident); its outputs are "
                "wrong")
_PAIR_LAST: dict = {}      # device index -> workspace of the last pair-form call
# device index -> int32 (1,): OR of the status words of every pair-form call
# since the last reset (a fill kernel clears each call's own word, so this is
# what survives later calls).  Made by the first pair-form call on the device;
# if that call is inside a graph capture, the graph also holds its zero fill
# and each replay starts the record afresh (an eager call first avoids that).
_PAIR_STICKY: dict = {}


def pair_applicable(degree: int, l: int, batch: int, device=None) -> bool:
    """Whether the pair form takes this shape: degree 1024 or 2048, l <= 3 and
    every workgroup resident at once (two workgroups per ciphertext, batch
    rounded up to 8; one workgroup per CU at N = 2048, two at N = 1024)."""
    if degree not in (1024, 2048) or not 1 <= l <= 3 or batch < 1:
        return False
    cus = torch.cuda.get_device_properties(device).multi_processor_count
    return (batch + 7) // 8 * 8 <= (cus if degree == 1024 else cus // 2)


def pair_status(device=None) -> int:
    """Status word of the last pair-form call on `device` (synchronises):
    bit 0 = a workgroup's wait on its partner ran out (that call's outputs are
    wrong), bit 1 = some pair ran on two XCDs; 0 when no call was made."""
    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    ws = _PAIR_LAST.get(dev)
    return 0 if ws is None else int(ws[:4].view(torch.int32).item())


def pair_errors(device=None, reset: bool = True) -> int:
    """OR of the status words of every pair-form call on `device` since the
    last reset (synchronises the device): bit 0 set means some call's outputs
    are wrong.  With `reset` the record starts again from zero.  Calls on one
    stream, or on streams ordered by events, add to it in order; two pair-form
    calls running at once on unordered streams may race on the word."""
    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    st = _PAIR_STICKY.get(dev)
    if st is None:
        return 0
    torch.cuda.synchronize(dev)
    word = int(st.item())
    if reset:
        st.zero_()
    return word


def pair_error_word(device=None):
    """The device tensor behind pair_errors (None before the first pair-form
    call on `device`), for callers that copy it to the host asynchronously."""
    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    return _PAIR_STICKY.get(dev)


def check_pair_errors(device=None) -> None:
    """Raises TFHEInternalError if a pair-form call on `device` since the last
    check had a partner wait run out (its outputs are wrong), and starts the
    record again.  Synchronises with the device."""
    if pair_errors(device, reset=True) & 1:
        raise _core.TFHEInternalError(PAIR_RUN_OUT)


def _pair_call(ct, bsk, arr, rot, idx, out, batch, log_beta, log_q, l, degree, slots=None,
               slot_count=0, spin_limit=None):
    ws_bytes = int(_core.lib().bslstm_mul_bootstrap_pair_workspace_bytes(batch, l, degree))
    if ws_bytes < 0:
        raise _core.TFHEInvalidArgument(
            f"the pair form runs at degree 1024 or 2048 and l <= 3 (got N={degree}, l={l})")
    ws = torch.empty(max(ws_bytes, 4), dtype=torch.int8, device=ct.device)
    spin = PAIR_SPIN_LIMIT if spin_limit is None else int(spin_limit)
    if slots is None:
        rot_rows, rot_cols = (0, 0) if rot is None else (rot.shape[-2], rot.shape[-1])
        _core.check(_core.lib().bslstm_mul_bootstrap_pair(
            _core.stream_of(ct), ct.data_ptr(), ct.numel(), bsk.data_ptr(), arr, _core.ptr(rot),
            rot_rows, rot_cols, idx.data_ptr(), idx.numel(), out.data_ptr(), batch, log_beta,
            log_q, l, degree, ws.data_ptr(), ws_bytes, spin))
    else:
        _core.check(_core.lib().bslstm_mul_bootstrap_pair_slots(
            _core.stream_of(ct), ct.data_ptr(), ct.numel(), bsk.data_ptr(), arr, slots.data_ptr(),
            slot_count, idx.data_ptr(), idx.numel(), out.data_ptr(), batch, log_beta, log_q, l,
            degree, ws.data_ptr(), ws_bytes, spin))
    dev = ct.device.index
    _PAIR_LAST[dev] = ws
    st = _PAIR_STICKY.get(dev)
    if st is None:
        st = _PAIR_STICKY[dev] = torch.zeros(1, dtype=torch.int32, device=ct.device)
    st.bitwise_or_(ws[:4].view(torch.int32))


SPLIT_CHUNK = 512      # ciphertexts per chunk of the split form's step scratch


def prepare_split_key(bsk_reshaped: torch.Tensor, l: int, degree: int) -> torch.Tensor:
    """FFT-domain key (n_iter, 2, l, 2, m, N/2) complex128 -> the split
    form's int8 Toeplitz limb rows (bslstm_mul_bootstrap_split_prepare_key).

    The result is cached on the key tensor itself (an attribute that lives
    and dies with it, tagged with the tensor's version so that a key written
    in place is prepared again); a freed key's memory reused by a new tensor
    can therefore never return a stale preparation."""
    _core.require_device(bsk_reshaped)
    bsk = bsk_reshaped.contiguous()
    dims = tuple(int(d) for d in bsk.shape[-6:])
    tag = (dims, bsk._version, l, degree)
    hit = getattr(bsk_reshaped, "_bslstm_split_key", None)
    if hit is not None and hit[0] == tag:
        return hit[1]
    arr = (_core._I64 * 6)(*dims)
    nbytes = int(_core.lib().bslstm_mul_bootstrap_split_key_bytes(arr, degree))
    if nbytes < 0:
        raise _core.TFHEInvalidArgument(f"the split form does not take this key {dims} at N={degree}")
    key = torch.empty(nbytes, dtype=torch.int8, device=bsk.device)
    _core.check(_core.lib().bslstm_mul_bootstrap_split_prepare_key(
        _core.stream_of(bsk), bsk.data_ptr(), arr, l, degree, key.data_ptr()))
    bsk_reshaped._bslstm_split_key = (tag, key)
    return key


def mul_bootstrap_split(ct_init: torch.Tensor, split_key: torch.Tensor, bsk_dims, indices: torch.Tensor,
                        log_beta: int, log_q: int, l: int, chunk: int = SPLIT_CHUNK) -> torch.Tensor:
    """Exact blind rotation on the matrix cores (bslstm_mul_bootstrap_split):
    same operands and results as mul_bootstrap with the key prepared by
    prepare_split_key; equals the int128 product bit for bit (log_q <= 32)."""
    _core.require_device(ct_init, split_key, indices)
    if ct_init.dtype != torch.int64 or indices.dtype != torch.int32:
        raise _core.TFHEInvalidArgument("ct_init must be int64 and indices int32")
    degree = ct_init.shape[-1]
    ct = ct_init.contiguous()
    idx = indices.contiguous()
    dims = [int(d) for d in bsk_dims]
    n_iter, m = dims[0], dims[4]
    batch = idx.numel() // (n_iter * m) if n_iter * m else 0
    out = torch.empty(batch, 2, degree, dtype=torch.int64, device=ct.device)
    arr = (_core._I64 * 6)(*dims)
    ws_bytes = int(_core.lib().bslstm_mul_bootstrap_split_workspace_bytes(batch, arr, degree, chunk))
    if ws_bytes < 0:
        raise _core.TFHEInvalidArgument(f"the split form does not take key dims {dims} at N={degree}")
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.int8, device=ct.device)
    _core.check(_core.lib().bslstm_mul_bootstrap_split(
        _core.stream_of(ct), ct.data_ptr(), ct.numel(), split_key.data_ptr(), arr, idx.data_ptr(),
        idx.numel(), out.data_ptr(), batch, log_beta, log_q, l, degree, ws.data_ptr(), ws_bytes))
    return out


_FORM_CODE = {"plain": 1, "batched": 2, "rows": 3}

# Launch-shape autotuner: the fastest form per (degree, l, n_iter, m, log2
# batch bucket, key slots, candidates), measured on the first live call of a
# shape (every candidate runs twice, the faster call counting; the forms give
# bit-identical results, so the call returns the last one's output).
_TUNED: dict = {}


def _bucket(batch: int) -> int:
    return max(0, int(batch - 1).bit_length())


def tuned_forms() -> dict:
    """The autotuner's choices so far: {(degree, l, n_iter, m, bucket, slots,
    candidates): form}."""
    return dict(_TUNED)


def save_tuning(path: str) -> None:
    """Write the autotuner's table to a JSON file (one record per shape)."""
    import json
    recs = [{"key": [k[0], k[1], k[2], k[3], k[4], bool(k[5]), list(k[6])], "form": f}
            for k, f in _TUNED.items()]
    with open(path, "w") as fh:
        json.dump({"format": "bslstm-tuning", "version": 1, "records": recs}, fh)


def load_tuning(path: str) -> int:
    """Merge a table written by save_tuning; returns the number of records.
    Records naming an unknown form are skipped."""
    import json
    with open(path) as fh:
        d = json.load(fh)
    if d.get("format") != "bslstm-tuning" or d.get("version") != 1:
        raise ValueError(f"{path} is not a version-1 tuning table")
    n = 0
    for r in d["records"]:
        k = r["key"]
        if r["form"] in FORMS:
            _TUNED[(int(k[0]), int(k[1]), int(k[2]), int(k[3]), int(k[4]), bool(k[5]),
                    tuple(k[6]))] = r["form"]
            n += 1
    return n


def mul_bootstrap(ct_init: torch.Tensor, bsk_reshaped: torch.Tensor, all_rot_fft: torch.Tensor | None,
                  indices: torch.Tensor, log_beta: int, log_q: int, l: int,
                  form: str = "batched", slots: torch.Tensor | None = None,
                  candidates: tuple = FORMS, spin_limit: int | None = None) -> torch.Tensor:
    """The raw custom call of cuTFHE/multiply.py:17-51 (handler
    csrc/mul_bootstrap.cuh:192-243): int64 ct (…,2,N) -> int64 (batch,2,N).

    form "plain": one workgroup per ciphertext (the reference's launch
    shape); "batched": two ciphertexts per workgroup sharing every key load;
    "rows": one ciphertext per workgroup with its two accumulator rows on two
    thread groups (the latency form for small batches);
    "pair": two workgroups per ciphertext, one accumulator row each,
    exchanging the digit spectra once per step (degree 1024 or 2048, l <= 3,
    batch up to CUs/2 at N = 2048 and CUs at N = 1024; the lower-latency form
    for up to a few hundred ciphertexts; spin_limit bounds each wait on the
    partner, see pair_status);
    "split": exact integer products as int8 MFMA GEMMs (log_q <= 32; the key
    is prepared once per key tensor and cached), equal to the exact product;
    "auto": the autotuner's choice for this shape (timed on first use; an
    untuned shape inside a HIP-graph capture is an error: capture after one
    eager call, or after load_tuning).  All forms give bit-identical results.

    slots (int32, one per ciphertext): bsk_reshaped is then a pool
    (C, n_iter, 2, l, 2, m, N/2) of client keys and ciphertext i uses
    pool[slots[i]]; in the batched form consecutive pairs must share a slot
    (callers whose pairs may not share one tune over candidates=("plain",
    "rows", "pair")).  Candidates that do not take the shape ("pair" outside
    its range) are left out of the tuning."""
    if form == "auto":
        dims = tuple(int(d) for d in bsk_reshaped.shape[-6:])
        batch = indices.numel() // (dims[0] * dims[4]) if dims[0] * dims[4] else 0
        if "pair" in candidates and not pair_applicable(ct_init.shape[-1], l, batch,
                                                        ct_init.device):
            candidates = tuple(c for c in candidates if c != "pair") or ("rows",)
        key = (ct_init.shape[-1], l, dims[0], dims[4], _bucket(batch), slots is not None,
               tuple(candidates))
        if key in _TUNED:
            form = _TUNED[key]
        elif batch == 0:
            form = candidates[0]
        elif torch.cuda.is_current_stream_capturing():
            # timing needs a synchronisation, which a capture cannot hold:
            # never record an untuned guess into a graph
            raise _core.TFHEInvalidArgument(
                f"form='auto' for an untuned shape {key[:5]} inside a graph capture; run the "
                f"same call once eagerly (or load_tuning) before capturing")
        else:
            # best of two calls per candidate: a candidate's first call may
            # allocate (its output, a workspace), which the first timing sees
            times = {}
            out = None
            for f in candidates:
                for _ in range(2):
                    a = torch.cuda.Event(enable_timing=True)
                    b = torch.cuda.Event(enable_timing=True)
                    a.record()
                    out = mul_bootstrap(ct_init, bsk_reshaped, all_rot_fft, indices, log_beta,
                                        log_q, l, f, slots)
                    b.record()
                    b.synchronize()
                    times[f] = min(times.get(f, float("inf")), a.elapsed_time(b))
            _TUNED[key] = min(times, key=times.get)
            return out
    if form == "split":
        if slots is not None:
            raise _core.TFHEInvalidArgument("the split form takes one key (no slots)")
        degree = ct_init.shape[-1]
        key = prepare_split_key(bsk_reshaped, l, degree)
        return mul_bootstrap_split(ct_init, key, bsk_reshaped.shape[-6:], indices, log_beta, log_q, l)
    if form not in FORMS:
        raise _core.TFHEInvalidArgument(f"form must be one of {FORMS}, 'split' or 'auto'")
    devs = [ct_init, bsk_reshaped, indices] + ([all_rot_fft] if all_rot_fft is not None else [])
    _core.require_device(*devs)
    if ct_init.dtype != torch.int64 or indices.dtype != torch.int32:
        raise _core.TFHEInvalidArgument("ct_init must be int64 and indices int32")
    if bsk_reshaped.dtype != torch.complex128 or bsk_reshaped.dim() < 6:
        raise _core.TFHEInvalidArgument("bsk must be complex128 with 6 trailing dims")
    degree = ct_init.shape[-1]
    ct = ct_init.contiguous()
    bsk = bsk_reshaped.contiguous()
    idx = indices.contiguous()
    dims = [int(d) for d in bsk.shape[-6:]]
    n_iter, m = dims[0], dims[4]
    batch = idx.numel() // (n_iter * m) if n_iter * m else 0
    out = torch.empty(batch, 2, degree, dtype=torch.int64, device=ct.device)
    arr = (_core._I64 * 6)(*dims)
    if slots is not None:
        _core.require_device(slots)
        sl = slots.to(torch.int32).reshape(-1).contiguous()
        if bsk.dim() != 7 or sl.numel() != batch:
            raise _core.TFHEInvalidArgument("slots need a (C, 6-D) key pool and one slot per ciphertext")
        if form == "pair":
            _pair_call(ct, bsk, arr, None, idx, out, batch, log_beta, log_q, l, degree, sl,
                       bsk.shape[0], spin_limit)
            return out
        _core.check(_core.lib().bslstm_mul_bootstrap_slots(
            _core.stream_of(ct), ct.data_ptr(), ct.numel(), bsk.data_ptr(), arr, sl.data_ptr(),
            bsk.shape[0], idx.data_ptr(), idx.numel(), out.data_ptr(), batch, log_beta, log_q, l,
            degree, _FORM_CODE[form]))
        return out
    rot = None if all_rot_fft is None else all_rot_fft.contiguous()
    if form == "pair":
        _pair_call(ct, bsk, arr, rot, idx, out, batch, log_beta, log_q, l, degree,
                   spin_limit=spin_limit)
        return out
    rot_rows, rot_cols = (0, 0) if rot is None else (rot.shape[-2], rot.shape[-1])
    entry = {"plain": _core.lib().bslstm_mul_bootstrap,
             "batched": _core.lib().bslstm_mul_bootstrap_batched,
             "rows": _core.lib().bslstm_mul_bootstrap_rows}[form]
    _core.check(entry(
        _core.stream_of(ct), ct.data_ptr(), ct.numel(), bsk.data_ptr(), arr,
        _core.ptr(rot), rot_rows, rot_cols, idx.data_ptr(), idx.numel(), out.data_ptr(), batch,
        log_beta, log_q, l, degree))
    return out


def multiply_seq_monomial(ciphertext, public_key: torch.Tensor, BSK: torch.Tensor, log_q: int,
                          log_beta: int, l: int, degree: int, collapse: int,
                          all_rot_fft: torch.Tensor | None, n_lut: int) -> torch.Tensor:
    """Blind rotation of a test polynomial by X^{sum a_i s_i}
    (cuTFHE/multiply.py:54-92).

    ciphertext   (…, 2, N) (or a (A, B) pair) test polynomial(s); one per
                 public key or a single one broadcast to all
    public_key   (…, n) LWE masks
    BSK          collapsed FFT-domain key, reference layout (2,2,l,n_eq,N/2)
                 or kernel layout (n_iter,2,l,2,2^c,N/2)
    Returns float64 (…, 2, N) like the reference.
    """
    ct_init = torch.stack(tuple(ciphertext)) if isinstance(ciphertext, (tuple, list)) else ciphertext
    n = public_key.shape[-1]
    idx = rotation_indices(public_key, log_q, degree, collapse, n_lut)
    bsk = prepare_bsk(BSK, n, collapse)
    out = mul_bootstrap(ct_init.to(torch.int64), bsk, all_rot_fft, idx, log_beta, log_q, l)
    return out.reshape(*public_key.shape[:-1], 2, degree).to(torch.float64)


def log2_int(v: float) -> int:
    """int(math.log2(v) + 1e-6), the rounding the reference applies to q and
    beta before the call (schemas/RLWE_jax.py:292)."""
    return int(math.log2(v) + 1e-6)
