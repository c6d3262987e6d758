"""NumPy model of the split-form (int8 MFMA) blind rotation's arithmetic.

Test helper: it restates, on the CPU and in the same data layout, what the
kernels of csrc/split_br.hip compute, so that the CPU suite can check the
method (Nussbaumer transform, limb cuts, Toeplitz rows, limb-pair sums, the
exact division) against the int128 oracle before any GPU run.

The exact negacyclic product in Z[X]/(X^N+1), N = M*R, R | M:
  a(X) = sum_{u<R} X^u A_u(Y),  Y = X^R,  A_u in Z[Y]/(Y^M+1)  (A_u[v] = a[u + R v])
  alpha = (A_0 .. A_{R-1}, 0 .. 0)   (length 2R)
  w = Y^(M/R) has order 2R in Z[Y]/(Y^M+1); a length-2R DFT over that ring
  only shifts and negates coefficients.  The product's length-2R cyclic
  convolution, inverted without the 1/2R, is 2R * gamma; folding with
  X^R = Y and dividing by 2R exactly gives the product.  All of it is
  Z-linear with integer coefficients, so it can run mod 2^(32 + log2(2R)) and
  the division leaves the product mod 2^32.
"""
from __future__ import annotations

import numpy as np

LIMBS = 5            # signed byte limbs of the transformed key (40 >= 32 + log2(2R) bits)


def ring_shape(N: int):
    """(M, R) with N = M*R and R | M, as the kernels choose."""
    return {512: (32, 16), 1024: (32, 32), 2048: (64, 32), 4096: (64, 64)}[N]


def mul_y(x, s: int, M: int):
    """x * Y^s in Z[Y]/(Y^M+1) along the last axis (s any integer)."""
    s %= 2 * M
    neg = s >= M
    s %= M
    out = np.roll(x, s, axis=-1)
    out[..., :s] = -out[..., :s]
    return -out if neg else out


def forward(poly, M: int, R: int):
    """Radix-2 decimation in frequency over the ring, natural-order input:
    (..., N) -> (..., 2R, M) with bins at their bit-reversed positions."""
    poly = np.asarray(poly, np.int64)
    lead = poly.shape[:-1]
    A = poly.reshape(*lead, M, R).swapaxes(-1, -2)          # A[u][v] = a[u + R v]
    x = np.zeros(lead + (2 * R, M), np.int64)
    x[..., :R, :] = A
    L = 2 * R
    # first stage: the upper half is zero
    for k in range(R):
        x[..., k + R, :] = mul_y(x[..., k, :], k * (M // R), M)
    half = R // 2
    while half >= 1:
        step = (L // (2 * half)) * (M // R)        # twiddle w^(k*L/(2 half)) = Y^(k*step)
        for b in range(0, L, 2 * half):
            for k in range(half):
                p, q = x[..., b + k, :].copy(), x[..., b + k + half, :].copy()
                x[..., b + k, :] = p + q
                x[..., b + k + half, :] = mul_y(p - q, k * step, M)
        half //= 2
    return x


def inverse_fold(bins, M: int, R: int):
    """Decimation in time on bit-reversed input with w^-1 (no 1/2R), then the
    fold C_u = g_u + Y g_(u+R): (..., 2R, M) -> (..., N), the product times 2R."""
    with np.errstate(over="ignore"):
        x = np.array(bins, np.int64, copy=True)
        L = 2 * R
        half = 1
        while half < L:
            step = (L // (2 * half)) * (M // R)
            for b in range(0, L, 2 * half):
                for k in range(half):
                    p = x[..., b + k, :].copy()
                    q = mul_y(x[..., b + k + half, :], -k * step, M)
                    x[..., b + k, :] = p + q
                    x[..., b + k + half, :] = p - q
            half *= 2
        C = x[..., :R, :] + mul_y(x[..., R:, :], 1, M)          # (..., R, M)
        return C.swapaxes(-1, -2).reshape(*C.shape[:-2], M * R)


def byte_limbs(v, nbits: int = 8 * LIMBS):
    """Balanced signed byte limbs of v mod 2^nbits: (..., LIMBS) int64 in
    [-128, 127] with sum limb_t 2^(8t) = v mod 2^(8*LIMBS)."""
    v = np.asarray(v, np.int64).astype(np.uint64) & np.uint64((1 << nbits) - 1)
    out = np.empty(v.shape + (LIMBS,), np.int64)
    carry = np.zeros(v.shape, np.int64)
    for t in range(LIMBS):
        d = ((v >> np.uint64(8 * t)) & np.uint64(255)).astype(np.int64) + carry
        carry = (d >= 128).astype(np.int64)
        out[..., t] = d - 256 * carry
    return out


def rev_vectors(kbin, M: int):
    """Key bins (..., M) -> the Toeplitz source rows (..., LIMBS, 2M) int8
    values: rev[x] = ext[M-1-x], ext[d] = K[d] (d >= 0), -K[d+M] (d < 0),
    each entry as limbs of its value mod 2^40; rev[2M-1] = 0."""
    d = M - 1 - np.arange(2 * M - 1)
    ext = np.where(d >= 0, kbin[..., np.clip(d, 0, M - 1)], -kbin[..., np.clip(d + M, 0, M - 1)])
    limbs = byte_limbs(ext)                                  # (..., 2M-1, LIMBS)
    out = np.zeros(kbin.shape[:-1] + (LIMBS, 2 * M), np.int64)
    out[..., :2 * M - 1] = np.moveaxis(limbs, -1, -2)
    return out


def toeplitz_rows(rev, M: int):
    """(..., LIMBS, 2M) -> (..., LIMBS, M[k_out], M[i]): entry rev[M-1-k_out+i],
    the negacyclic Toeplitz matrix of the key bin (the kernel reads 16 of these
    per lane from x0 = M-1-k_out+i0)."""
    k = np.arange(M)[:, None]
    i = np.arange(M)[None, :]
    return rev[..., M - 1 - k + i]


def decompose(ct, log_beta: int, log_q: int, L: int):
    """Offset-rule digits (csrc/dec_fft.cuh:4-31): (2, N) -> (L, 2, N)."""
    off = 0
    for i in range(1, L + 1):
        off += 1 << (log_q - i * log_beta + log_beta - 1)
    if log_q - L * log_beta - 1 >= 0:
        off += 1 << (log_q - L * log_beta - 1)
    st = (np.asarray(ct, np.int64).astype(np.uint64) + np.uint64(off))
    out = np.empty((L,) + st.shape, np.int64)
    for i in range(1, L + 1):
        sh = log_q - i * log_beta
        out[i - 1] = ((st >> np.uint64(sh)) & np.uint64((1 << log_beta) - 1)).astype(np.int64) \
            - (1 << (log_beta - 1))
    return out


def rotate(poly, e: int):
    """X^e * poly in Z[X]/(X^N+1), e in [0, 2N)."""
    N = poly.shape[-1]
    e %= 2 * N
    neg = e >= N
    e %= N
    out = np.roll(poly, e, axis=-1)
    out[..., :e] = -out[..., :e]
    return -out if neg else out


def prepare_key(bsk_coef_k, N: int):
    """Coefficient key (n_iter, 2[a], L, 2[comp], m, N) -> Toeplitz sources
    (n_iter, 2R[bin], P[pbc], 2[a], LIMBS, 2M), pbc = (2*lev + comp)*m + p."""
    M, R = ring_shape(N)
    k = np.asarray(bsk_coef_k, np.int64)
    n_iter, _, L, _, m, _ = k.shape
    k = k.reshape(n_iter, 2, 2 * L * m, N)                   # (g, a, pbc, N)
    kb = forward(k, M, R)                                     # (g, a, pbc, 2R, M)
    kb = np.transpose(kb, (0, 3, 2, 1, 4))                    # (g, bin, pbc, a, M)
    return rev_vectors(kb, M)


def external_product(acc, key_g, idx_g, log_beta: int, log_q: int, L: int):
    """One step for one ciphertext: acc (2, N) classes, key_g (2R, P, 2,
    LIMBS, 2M), idx_g (m,) -> new acc (2, N) centred mod q."""
    N = acc.shape[-1]
    M, R = ring_shape(N)
    m = len(idx_g)
    dig = decompose(acc, log_beta, log_q, L).reshape(2 * L, N)        # bc = 2*lev + comp
    rot = np.stack([rotate(dig[bc], int(idx_g[p])) for bc in range(2 * L) for p in range(m)])
    db = forward(rot, M, R)                                   # (P, 2R, M), |.| <= R*beta/2
    assert np.abs(db).max(initial=0) <= R << (log_beta - 1)
    lo = ((db + 128) & 255) - 128
    hi = (db - lo) >> 8
    dlimb = (lo, hi)
    T = toeplitz_rows(key_g, M)                               # (2R, P, 2, LIMBS, M, M)
    out = np.zeros((2, 2 * R, M), np.int64)
    for u in range(LIMBS):
        acc_u = np.zeros((2, 2 * R, M), np.int64)
        for s in (0, 1):
            t = u - s
            if not 0 <= t < LIMBS:
                continue
            # C[a, bin, k_out] = sum_{pbc, i} T[bin, pbc, a, t, k_out, i] * D_s[pbc, bin, i]
            acc_u += np.einsum("jpakI,pjI->ajk", T[:, :, :, t], dlimb[s])
        assert np.abs(acc_u).max(initial=0) < 2 ** 31          # int32 MFMA accumulators
        with np.errstate(over="ignore"):
            out += acc_u << (8 * u)
    with np.errstate(over="ignore"):
        prod2r = inverse_fold(out, M, R)                      # 2R * product, mod 2^64
    lg = (2 * R).bit_length() - 1
    c = (prod2r.astype(np.uint64) >> np.uint64(lg)) & np.uint64(0xFFFFFFFF)
    c = c & np.uint64((1 << log_q) - 1)
    c = c.astype(np.int64)
    return np.where(c >= (1 << (log_q - 1)), c - (1 << log_q), c)


def blind_rotate(ct, bsk_coef_k, idx, log_beta: int, log_q: int, L: int):
    """Split-form blind rotation model: (B,2,N)/(2,N), (n_iter,2,L,2,m,N),
    (B,n_iter,m) -> (B,2,N) centred mod q (log_q <= 32)."""
    ct = np.asarray(ct, np.int64)
    idx = np.asarray(idx)
    B, n_iter, _ = idx.shape
    N = ct.shape[-1]
    key = prepare_key(bsk_coef_k, N)
    out = np.empty((B, 2, N), np.int64)
    for b in range(B):
        acc = ct if ct.ndim == 2 else ct[b]
        for g in range(n_iter):
            acc = external_product(acc, key[g], idx[b, g], log_beta, log_q, L)
        out[b] = acc
    return out
