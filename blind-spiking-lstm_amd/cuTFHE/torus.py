"""Device entry points of the LSTM path for int64 torus values.

Thin, validating wrappers over the C ABI (include/bslstm_tfhe.h): every
function takes and returns torch tensors on the GPU and runs one or more
hand-written HIP kernels on the tensor's current stream.  Torus values are
int64 classes mod q = 2^log_q (log_q <= 64) held as centred representatives;
complex128 tensors hold polynomials in the reference's transform domain
(F_j = p(zeta^(2j+1)), j < N/2).
"""
from __future__ import annotations

import torch

from . import _core


def _i64(x: torch.Tensor) -> torch.Tensor:
    """int64 view of a torus tensor: int64 as is; float64 (the reference's
    carrier, exact below 2^53) rounded to the nearest integer."""
    if x.dtype == torch.int64:
        return x.contiguous()
    if x.dtype in (torch.float64, torch.float32):
        return torch.round(x.to(torch.float64)).to(torch.int64).contiguous()
    return x.to(torch.int64).contiguous()


def _lib():
    return _core.lib()


def poly_fft(polys: torch.Tensor) -> torch.Tensor:
    """jax_fourier (polynomial_jax.py:11-15) of integer polynomials on the
    GPU: (..., N) int64 -> (..., N/2) complex128."""
    _core.require_device(polys)
    p = _i64(polys)
    N = p.shape[-1]
    cnt = p.numel() // N if N else 0
    out = torch.empty(*p.shape[:-1], N // 2, dtype=torch.complex128, device=p.device)
    _core.check(_lib().bslstm_poly_fft(_core.stream_of(p), p.data_ptr(), cnt, N,
                                       out.data_ptr()))
    return out


def poly_ifft(spectra: torch.Tensor) -> torch.Tensor:
    """Inverse of poly_fft (jax_inversefourier, polynomial_jax.py:17-21):
    (..., N/2) complex128 -> (..., N) int64, rounded to the nearest integer."""
    _core.require_device(spectra)
    sp = spectra.to(torch.complex128).contiguous()
    N = 2 * sp.shape[-1]
    cnt = sp.numel() // (N // 2) if N else 0
    out = torch.empty(*sp.shape[:-1], N, dtype=torch.int64, device=sp.device)
    _core.check(_lib().bslstm_poly_ifft(_core.stream_of(sp), sp.data_ptr(), cnt, N,
                                        out.data_ptr()))
    return out


def linear_extract(ct: torch.Tensor, weights: torch.Tensor, bias: torch.Tensor | None,
                   log_q: int, support: torch.Tensor | None = None, check_support: bool = True):
    """Homomorphic Linear + coefficient-0 extraction (Spike_LSTM.py:46-61).

    ct (batch, 2, N) or (2, N); weights (out, N) int64 weight polynomials;
    bias (out,) or None.  Returns (masks (batch, out, N), bodies (batch, out))
    (the batch axis is dropped for a (2, N) input)."""
    _core.require_device(ct, weights)
    c = _i64(ct)
    w = _i64(weights)
    N = c.shape[-1]
    single = c.dim() == 2
    batch = 1 if single else c.numel() // (2 * N)
    out_dim = w.shape[0]
    if w.shape[-1] != N:
        raise _core.TFHEInvalidArgument(f"weights have {w.shape[-1]} columns, ciphertext N={N}")
    if support is None:
        support = torch.nonzero(torch.any(w != 0, dim=0)).reshape(-1).to(torch.int32)
    support = support.to(device=c.device, dtype=torch.int32).contiguous()
    # (callers that built `support` from the weights themselves skip this
    # device->host check, e.g. inside a captured graph)
    if check_support and support.numel() and (int(support.min()) < 0 or int(support.max()) >= N):
        raise _core.TFHEInvalidArgument("support entries must lie in [0, N)")
    b = None if bias is None else _i64(bias.reshape(-1))
    masks = torch.empty(batch, out_dim, N, dtype=torch.int64, device=c.device)
    bodies = torch.empty(batch, out_dim, dtype=torch.int64, device=c.device)
    _core.check(_lib().bslstm_linear_extract(
        _core.stream_of(c), c.data_ptr(), c.numel(), w.data_ptr(),
        support.data_ptr() if support.numel() else None, support.numel(),
        None if b is None else b.data_ptr(), batch, out_dim, N, log_q, masks.data_ptr(),
        bodies.data_ptr()))
    if single:
        return masks[0], bodies[0]
    return masks, bodies


def pack_lwe_to_rlwe(a: torch.Tensor, b: torch.Tensor, galois_key_fft: torch.Tensor,
                     log_beta: int, l: int, log_q: int) -> torch.Tensor:
    """Trace packing (text_jax.py:314-345) of LWEs a (..., N), b (...) into
    RLWEs (..., 2, N)."""
    _core.require_device(a, b, galois_key_fft)
    ai = _i64(a)
    N = ai.shape[-1]
    cnt = ai.numel() // N
    bi = _i64(b).reshape(-1)
    if bi.numel() != cnt:
        raise _core.TFHEInvalidArgument(f"{bi.numel()} bodies for {cnt} masks")
    gk = galois_key_fft.to(torch.complex128).contiguous()
    logn = N.bit_length() - 1
    if tuple(gk.shape) != (logn, l, 2, N // 2):
        raise _core.TFHEInvalidArgument(
            f"galois key shape {tuple(gk.shape)}; expected {(logn, l, 2, N // 2)}")
    out = torch.empty(*ai.shape[:-1], 2, N, dtype=torch.int64, device=ai.device)
    _core.check(_lib().bslstm_pack_lwe_to_rlwe(_core.stream_of(ai), ai.data_ptr(), bi.data_ptr(),
                                               cnt, gk.data_ptr(), N, log_beta, l, log_q,
                                               out.data_ptr()))
    return out


def lut_product_sum(rlwe: torch.Tensor, lut_fft: torch.Tensor, n_lut: int,
                    log_q: int) -> torch.Tensor:
    """sum_k round(LUT[k, g] * RLWE[s, k, g]) (activation.py:103-106 summed as
    in Spike_LSTM.py:129-132).  rlwe (S, n_lut, G, 2, N) or (n_lut, G, 2, N),
    lut_fft (n_lut, G, N/2) shared over S -> (S, G, 2, N) or (G, 2, N)."""
    _core.require_device(rlwe, lut_fft)
    r = _i64(rlwe)
    N = r.shape[-1]
    lf = lut_fft.to(torch.complex128).contiguous()
    G = lf.numel() // (n_lut * (N // 2))
    if G == 0 or lf.numel() != n_lut * G * (N // 2) or r.numel() % (n_lut * G * 2 * N):
        raise _core.TFHEInvalidArgument("rlwe/lut shapes do not match (S, n_lut, G, 2, N)")
    S = r.numel() // (n_lut * G * 2 * N)
    out = torch.empty(S, G, 2, N, dtype=torch.int64, device=r.device)
    _core.check(_lib().bslstm_lut_product_sum(_core.stream_of(r), r.data_ptr(), lf.data_ptr(), S,
                                              G, n_lut, N, log_q, out.data_ptr()))
    return out if r.dim() == 5 else out[0]


def packing(masks: torch.Tensor, bodies: torch.Tensor, ksk_fft: torch.Tensor, log_beta: int,
            l: int, log_q: int) -> torch.Tensor:
    """packing (text_jax.py:211-239): masks (P, n_lwes, n_in), bodies
    (P, n_lwes) -> RLWE (P, 2, N); ksk_fft (n_in, l, 2, N/2)."""
    _core.require_device(masks, bodies, ksk_fft)
    m = _i64(masks)
    if m.dim() == 2:
        m = m.unsqueeze(0)
    P, n_lwes, n_in = m.shape
    bd = _i64(bodies).reshape(P, n_lwes)
    kf = ksk_fft.to(torch.complex128).contiguous()
    if kf.dim() != 4 or kf.shape[0] != n_in or kf.shape[1] != l or kf.shape[2] != 2:
        raise _core.TFHEInvalidArgument(
            f"packing key shape {tuple(kf.shape)}; expected ({n_in}, {l}, 2, N/2)")
    N = 2 * kf.shape[3]
    ws_bytes = int(_lib().bslstm_packing_workspace_bytes(P, n_in, N))
    if ws_bytes < 0:
        raise _core.TFHEInvalidArgument("bad packing shapes")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=m.device)
    out = torch.empty(P, 2, N, dtype=torch.int64, device=m.device)
    _core.check(_lib().bslstm_packing(_core.stream_of(m), m.data_ptr(), bd.data_ptr(), P, n_lwes,
                                      n_in, kf.data_ptr(), N, log_beta, l, log_q, ws.data_ptr(),
                                      ws_bytes, out.data_ptr()))
    return out


def axpy(x: torch.Tensor, y: torch.Tensor, sx: int, sy: int, log_q: int) -> torch.Tensor:
    """centred(sx*x + sy*y) mod 2^log_q, elementwise on int64 tensors."""
    _core.require_device(x, y)
    xi, yi = _i64(x), _i64(y)
    if xi.shape != yi.shape:
        xi, yi = torch.broadcast_tensors(xi, yi)
        xi, yi = xi.contiguous(), yi.contiguous()
    out = torch.empty_like(xi)
    _core.check(_lib().bslstm_torus_axpy(_core.stream_of(xi), xi.data_ptr(), yi.data_ptr(),
                                         xi.numel(), sx, sy, log_q, out.data_ptr()))
    return out


def rotate_poly(x: torch.Tensor, exponents, log_q: int) -> torch.Tensor:
    """X^e * x for x (..., N); exponents an int, or a tensor with one
    exponent per leading index of x (rows sharing an exponent are
    consecutive)."""
    _core.require_device(x)
    xi = _i64(x)
    N = xi.shape[-1]
    rows = xi.numel() // N
    if isinstance(exponents, torch.Tensor) and exponents.device == xi.device:
        e = exponents.to(torch.int64).reshape(-1).contiguous()
    else:
        e = torch.as_tensor(exponents, dtype=torch.int64).reshape(-1).to(xi.device).contiguous()
    if rows % e.numel():
        raise _core.TFHEInvalidArgument(f"{e.numel()} exponents for {rows} rows")
    out = torch.empty_like(xi)
    _core.check(_lib().bslstm_rotate_poly(_core.stream_of(xi), xi.data_ptr(), rows, N,
                                          e.data_ptr(), rows // e.numel(), log_q, out.data_ptr()))
    return out


def key_switch(ksk_a: torch.Tensor, ksk_b: torch.Tensor, a: torch.Tensor, b: torch.Tensor,
               log_beta: int, l: int, log_q: int):
    """Exact LWE key switch (LWE_jax.py:97-117) on int64: a (B, n_in), b (B,)
    -> (a' (B, n_out), b' (B,))."""
    _core.require_device(ksk_a, ksk_b, a, b)
    ai = _i64(a)
    n_in = ai.shape[-1]
    batch = ai.numel() // n_in
    bi = _i64(b).reshape(-1)
    ka, kb = _i64(ksk_a), _i64(ksk_b)
    n_out = ka.shape[-1]
    if ka.numel() != l * n_in * n_out or kb.numel() != l * n_in or bi.numel() != batch:
        raise _core.TFHEInvalidArgument("key-switch key / ciphertext shapes do not match")
    a_out = torch.empty(batch, n_out, dtype=torch.int64, device=ai.device)
    b_out = torch.empty(batch, dtype=torch.int64, device=ai.device)
    _core.check(_lib().bslstm_key_switch_lwe_i64(_core.stream_of(ai), ai.data_ptr(), bi.data_ptr(),
                                                 ka.data_ptr(), kb.data_ptr(), batch, n_in, n_out,
                                                 log_beta, l, log_q, a_out.data_ptr(),
                                                 b_out.data_ptr()))
    return a_out, b_out


def rotation_indices(public_key: torch.Tensor, log_q: int, degree: int, collapse: int,
                     n_lut: int = 1) -> torch.Tensor:
    """Rotation exponents (cuTFHE/multiply.py:67-79) from int64 masks."""
    _core.require_device(public_key)
    pk = _i64(public_key)
    n = pk.shape[-1]
    if n % collapse:
        raise _core.TFHEInvalidArgument(f"n={n} is not a multiple of collapse={collapse}")
    batch = pk.numel() // n if n else 0
    out = torch.empty(*pk.shape[:-1], n // collapse, 2 ** collapse, dtype=torch.int32,
                      device=pk.device)
    _core.check(_lib().bslstm_rotation_indices_i64(_core.stream_of(pk), pk.data_ptr(), batch, n,
                                                   collapse, log_q, degree, n_lut,
                                                   out.data_ptr()))
    return out


def rotate_extract(br_out: torch.Tensor, body: torch.Tensor | None, log_q: int, n_lut: int,
                   n_ext: int | None = None):
    """X^(-b_ms) rotation (none when body is None) + sample extraction of
    coefficients 0..n_ext-1: (B, 2, N) -> (masks (B, n_ext, N), bodies (B, n_ext))."""
    _core.require_device(br_out)
    br = _i64(br_out)
    N = br.shape[-1]
    batch = br.numel() // (2 * N)
    n_ext = n_lut if n_ext is None else n_ext
    bd = None if body is None else _i64(body).reshape(-1)
    if bd is not None and bd.numel() != batch:
        raise _core.TFHEInvalidArgument(f"{bd.numel()} bodies for {batch} ciphertexts")
    masks = torch.empty(batch, n_ext, N, dtype=torch.int64, device=br.device)
    bodies = torch.empty(batch, n_ext, dtype=torch.int64, device=br.device)
    for s in range(0, batch, 65535):
        e = min(batch, s + 65535)
        _core.check(_lib().bslstm_rotate_extract_i64(
            _core.stream_of(br), br[s:e].data_ptr(),
            None if bd is None else bd[s:e].data_ptr(), e - s, N, log_q, n_lut, n_ext,
            masks[s:e].data_ptr(), bodies[s:e].data_ptr()))
    return masks, bodies


def prepare_ks_key(ksk_a: torch.Tensor, ksk_b: torch.Tensor) -> torch.Tensor:
    """Key-switching key (l, n_in, n_out) / (l, n_in[, 1]) int64 -> the int8
    limb tiles of the MFMA key switch (bslstm_ks_mfma_prepare_key), once per
    key."""
    _core.require_device(ksk_a, ksk_b)
    ka, kb = _i64(ksk_a), _i64(ksk_b)
    l, n_in, n_out = ka.shape
    nbytes = int(_lib().bslstm_ks_mfma_key_bytes(l, n_in, n_out))
    if nbytes < 0:
        raise _core.TFHEInvalidArgument("key too large for the MFMA key switch")
    tiles = torch.empty(nbytes, dtype=torch.int8, device=ka.device)
    _core.check(_lib().bslstm_ks_mfma_prepare_key(_core.stream_of(ka), ka.data_ptr(), kb.data_ptr(),
                                                  l, n_in, n_out, tiles.data_ptr()))
    return tiles


def key_switch_mfma(key_tiles: torch.Tensor, a: torch.Tensor, b: torch.Tensor, n_out: int,
                    log_beta: int, l: int, log_q: int):
    """Exact LWE key switch on the matrix cores (int8 limb GEMMs):
    a (B, n_in), b (B,) -> (a' (B, n_out), b' (B,)), bit-identical to
    key_switch."""
    _core.require_device(key_tiles, a, b)
    ai = _i64(a)
    n_in = ai.shape[-1]
    batch = ai.numel() // n_in
    bi = _i64(b).reshape(-1)
    if bi.numel() != batch:
        raise _core.TFHEInvalidArgument("one body per mask expected")
    ws_bytes = int(_lib().bslstm_ks_mfma_workspace_bytes(batch, l, n_in))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.int8, device=ai.device)
    a_out = torch.empty(batch, n_out, dtype=torch.int64, device=ai.device)
    b_out = torch.empty(batch, dtype=torch.int64, device=ai.device)
    _core.check(_lib().bslstm_key_switch_lwe_mfma(
        _core.stream_of(ai), ai.data_ptr(), bi.data_ptr(), key_tiles.data_ptr(), batch, n_in, n_out,
        log_beta, l, log_q, ws.data_ptr(), ws_bytes, a_out.data_ptr(), b_out.data_ptr()))
    return a_out, b_out


def key_switch_mfma_noalloc(key_tiles: torch.Tensor, a: torch.Tensor, b: torch.Tensor,
                            n_out: int, log_beta: int, l: int, log_q: int, a_out: torch.Tensor,
                            b_out: torch.Tensor) -> None:
    """key_switch_mfma into caller-owned outputs a_out (B, n_out), b_out (B,)
    through the allocation-free entry point (digit tiles in the library's
    static per-device scratch; nothing is allocated by this call)."""
    _core.require_device(key_tiles, a, b, a_out, b_out)
    for t in (a, b, a_out, b_out):
        if t.dtype != torch.int64 or not t.is_contiguous():
            raise _core.TFHEInvalidArgument("contiguous int64 operands expected")
    n_in = a.shape[-1]
    batch = a.numel() // n_in
    if b.numel() != batch or a_out.numel() != batch * n_out or b_out.numel() != batch:
        raise _core.TFHEInvalidArgument("operand sizes do not match the batch")
    _core.check(_lib().bslstm_key_switch_lwe_mfma_noalloc(
        _core.stream_of(a), a.data_ptr(), b.data_ptr(), key_tiles.data_ptr(), batch, n_in, n_out,
        log_beta, l, log_q, a_out.data_ptr(), b_out.data_ptr()))


# ---------------------------------------------------------------------------
# key-slot forms: one launch over a batch that mixes clients, each item
# reading its client's key from a pool of equally shaped keys
# ---------------------------------------------------------------------------


def _slots(slots: torch.Tensor, count: int, pool_slots: int) -> torch.Tensor:
    sl = slots.to(torch.int32).reshape(-1).contiguous()
    if sl.numel() != count:
        raise _core.TFHEInvalidArgument(f"{sl.numel()} slots for {count} items")
    del pool_slots      # kernels clamp slot indices to the pool
    return sl


def pack_lwe_to_rlwe_slots(a, b, gk_pool: torch.Tensor, slots: torch.Tensor, log_beta: int,
                           l: int, log_q: int) -> torch.Tensor:
    """pack_lwe_to_rlwe with LWE i using the Galois key gk_pool[slots[i]]
    (gk_pool (C, log2 N, l, 2, N/2))."""
    _core.require_device(a, b, gk_pool, slots)
    ai = _i64(a)
    N = ai.shape[-1]
    cnt = ai.numel() // N
    bi = _i64(b).reshape(-1)
    gk = gk_pool.contiguous()
    logn = N.bit_length() - 1
    if tuple(gk.shape[1:]) != (logn, l, 2, N // 2) or bi.numel() != cnt:
        raise _core.TFHEInvalidArgument("galois key pool / LWE shapes do not match")
    sl = _slots(slots, cnt, gk.shape[0])
    out = torch.empty(*ai.shape[:-1], 2, N, dtype=torch.int64, device=ai.device)
    _core.check(_lib().bslstm_pack_lwe_to_rlwe_slots(
        _core.stream_of(ai), ai.data_ptr(), bi.data_ptr(), cnt, gk.data_ptr(), sl.data_ptr(),
        gk.shape[0], N, log_beta, l, log_q, out.data_ptr()))
    return out


def packing_slots(masks, bodies, ksp_pool: torch.Tensor, slots: torch.Tensor, log_beta: int,
                  l: int, log_q: int) -> torch.Tensor:
    """packing with packing p using ksp_pool[slots[p]] (ksp_pool (C, n_in, l, 2, N/2))."""
    _core.require_device(masks, bodies, ksp_pool, slots)
    m = _i64(masks)
    P, n_lwes, n_in = m.shape
    bd = _i64(bodies).reshape(P, n_lwes)
    kp = ksp_pool.contiguous()
    if kp.dim() != 5 or kp.shape[1] != n_in or kp.shape[2] != l:
        raise _core.TFHEInvalidArgument("packing key pool shape does not match")
    N = 2 * kp.shape[4]
    sl = _slots(slots, P, kp.shape[0])
    ws_bytes = int(_lib().bslstm_packing_workspace_bytes(P, n_in, N))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=m.device)
    out = torch.empty(P, 2, N, dtype=torch.int64, device=m.device)
    _core.check(_lib().bslstm_packing_slots(
        _core.stream_of(m), m.data_ptr(), bd.data_ptr(), P, n_lwes, n_in, kp.data_ptr(),
        sl.data_ptr(), kp.shape[0], N, log_beta, l, log_q, ws.data_ptr(), ws_bytes, out.data_ptr()))
    return out


def key_switch_mfma_slots(tiles_pool: torch.Tensor, tile_slots: torch.Tensor, a, b, n_out: int,
                          log_beta: int, l: int, log_q: int):
    """key_switch_mfma with row tile t (32 rows) using tiles_pool[tile_slots[t]]
    (tiles_pool (C, key bytes) int8)."""
    _core.require_device(tiles_pool, tile_slots, a, b)
    ai = _i64(a)
    n_in = ai.shape[-1]
    batch = ai.numel() // n_in
    bi = _i64(b).reshape(-1)
    sl = _slots(tile_slots, (batch + 31) // 32, tiles_pool.shape[0])
    ws_bytes = int(_lib().bslstm_ks_mfma_workspace_bytes(batch, l, n_in))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.int8, device=ai.device)
    a_out = torch.empty(batch, n_out, dtype=torch.int64, device=ai.device)
    b_out = torch.empty(batch, dtype=torch.int64, device=ai.device)
    _core.check(_lib().bslstm_key_switch_lwe_mfma_slots(
        _core.stream_of(ai), ai.data_ptr(), bi.data_ptr(), tiles_pool.data_ptr(), sl.data_ptr(),
        tiles_pool.shape[0], batch, n_in, n_out, log_beta, l, log_q, ws.data_ptr(), ws_bytes,
        a_out.data_ptr(), b_out.data_ptr()))
    return a_out, b_out
