"""CPU oracle for the TFHE programmable-bootstrapping path.

TEST INFRASTRUCTURE ONLY.  Nothing under ``blind-spiking-lstm_amd/`` may import,
call or link this module; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

It restates, in NumPy, the algorithm of the reference's bootstrapping path
(AMIAD-Research/Blind-Spiking-LSTM).  Each function cites the reference
file:line it follows.  Torus values are int64 integers (any representative of
their class mod q = 2**log_q, log_q <= 32 for the exact paths); the reference
carries the same integers in float64.

Parity pins (what checks each function; see tests/test_oracle.py):
  all_binary_vectors      closed form (bit j of row p is (p >> (c-1-j)) & 1)
  centered_mod            definition; brute force
  decompose_int           recomposition invariant, exhaustive on small q
  decompose_float         recomposition invariant; agrees with decompose_int
                          off ties
  negacyclic_fft/ifft     closed form (evaluation at zeta^(2j+1)); O(N^2)
                          direct evaluation; exact product via the C oracle
  rotation_indices        brute force with Python Fractions incl. built ties
  coef_rotation           X^N = -1 power series; composition laws
  sample_extract          LWE-decrypt(SE(ct,k)) == RLWE-decrypt(ct)[k], noise-free
  blind_rotate_f64        C exact int128 oracle (oracle/c/tfhe_exact.c);
                          decrypt == LUT(m) exhaustively over message space
  key_switch_lwe          exact integer identity; decrypt preserved
  get_plain_LUT           closed form of the interleave and X^(-N/2)
  FP64 FFT arithmetic of the reference (cuFFTDx, jnp.fft): parity unpinned at
  ULP level -- only the rounded integer results are compared.
"""
from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# small helpers
# ---------------------------------------------------------------------------


def all_binary_vectors(c: int) -> np.ndarray:
    """All 2^c binary patterns in meshgrid 'ij' order.

    Follows cuTFHE/utils.py:5-8 (and its twin schemas/text_jax.py:440-454):
    row p holds the bits of p, most significant first.
    """
    p = np.arange(2 ** c)[:, None]
    j = np.arange(c)[None, :]
    return ((p >> (c - 1 - j)) & 1).astype(np.int32)


def centered_mod(x, log_q: int):
    """x - q*rint(x/q) on integers (schemas/polynomial_jax.py:7-9).

    Returns the representative in [-q/2, q/2).  The reference's float version
    can also return +q/2; every comparison in the tests is made mod q.
    """
    q = 1 << log_q
    x = np.asarray(x, dtype=np.int64)
    r = np.mod(x, q)
    return np.where(r >= q // 2, r - q, r).astype(np.int64)


def eq_mod(a, b, log_q: int) -> bool:
    q = 1 << log_q
    return bool(np.all(np.mod(np.asarray(a, np.int64) - np.asarray(b, np.int64), q) == 0))


# ---------------------------------------------------------------------------
# gadget decompositions
# ---------------------------------------------------------------------------


def decompose_int(x, log_beta: int, log_q: int, L: int) -> np.ndarray:
    """Signed gadget digits by the integer offset rule of csrc/dec_fft.cuh:4-31.

    offset = sum_{i=1..L} 2^(log_q - i*log_beta + log_beta - 1)
             + 2^(log_q - L*log_beta - 1)            (dec_fft.cuh:7-17)
    digit_i = ((x + offset) >> (log_q - i*log_beta)) & (beta-1) - beta/2
                                                     (dec_fft.cuh:19-29)
    Output shape (L, *x.shape); digits lie in [-beta/2, beta/2).
    """
    x = np.asarray(x, dtype=np.int64)
    off = 0
    for i in range(1, L + 1):
        off += 1 << (log_q - i * log_beta + log_beta - 1)
    rs = log_q - L * log_beta - 1
    if 0 <= rs < 64:
        off += 1 << rs
    state = x.astype(np.uint64) + np.uint64(off)        # wraps mod 2^64
    out = np.empty((L,) + x.shape, dtype=np.int64)
    mask = np.uint64((1 << log_beta) - 1)
    for i in range(1, L + 1):
        sh = np.uint64(log_q - i * log_beta)
        out[i - 1] = ((state >> sh) & mask).astype(np.int64) - (1 << (log_beta - 1))
    return out


def decompose_float(gamma, beta: float, l: int, q: float) -> np.ndarray:
    """Float digit recursion of schemas/polynomial_jax.py:90-111.

    x0 = gamma/q; digit = round_half_even(x*beta); x <- x*beta - digit.
    Output (l, *gamma.shape) float64.
    """
    x = np.asarray(gamma, dtype=np.float64) / q
    out = []
    for _ in range(l):
        xb = x * beta
        d = np.rint(xb)
        x = xb - d
        out.append(d)
    return np.stack(out)


# ---------------------------------------------------------------------------
# negacyclic FFT in the reference convention
# ---------------------------------------------------------------------------


def zeta_pow(e, N: int) -> np.ndarray:
    """zeta^e with zeta = exp(-i*pi/N), e integer (taken mod 2N)."""
    e = np.mod(np.asarray(e, dtype=np.int64), 2 * N)
    # cos/sin of pi*e/N from the reduced integer keeps full accuracy
    ang = np.pi * e.astype(np.float64) / N
    return np.cos(ang) - 1j * np.sin(ang)


def negacyclic_fft(p) -> np.ndarray:
    """jax_fourier of schemas/polynomial_jax.py:11-15.

    F_j = sum_k p_k exp(-i*pi*k*(2j+1)/N) for j < N/2 (the same transform the
    reference GPU kernel builds in csrc/dec_fft.cuh:33-57).
    """
    p = np.asarray(p, dtype=np.float64)
    N = p.shape[-1]
    tw = zeta_pow(np.arange(N), N)
    return np.fft.fft(tw * p, axis=-1)[..., : N // 2]


def negacyclic_ifft_raw(F) -> np.ndarray:
    """Real coefficients before rounding: the inverse of polynomial_jax.py:17-21
    and of csrc/mul_bootstrap.cuh:70-90 (Hermitian extension, N-point IFFT,
    untwist by exp(+i*pi*k/N), real part)."""
    F = np.asarray(F)
    M = F.shape[-1]
    N = 2 * M
    full = np.concatenate([F, np.conj(F[..., ::-1])], axis=-1)
    tw = np.conj(zeta_pow(np.arange(N), N))
    return np.real(tw * np.fft.ifft(full, axis=-1))


def negacyclic_ifft(F) -> np.ndarray:
    """jax_inversefourier (polynomial_jax.py:17-21): rounded real coefficients."""
    return np.round(negacyclic_ifft_raw(F))


def negacyclic_mul_direct(a, b, log_q: int | None = None) -> np.ndarray:
    """Exact negacyclic product by the schoolbook rule with Python ints.

    Slow (O(N^2) Python); used on small N to pin the FFT and the C oracle.
    """
    a = [int(v) for v in np.asarray(a).ravel()]
    b = [int(v) for v in np.asarray(b).ravel()]
    N = len(a)
    out = [0] * N
    for i in range(N):
        ai = a[i]
        if ai == 0:
            continue
        for j in range(N):
            k = i + j
            if k < N:
                out[k] += ai * b[j]
            else:
                out[k - N] -= ai * b[j]
    if log_q is not None:
        q = 1 << log_q
        out = [((v % q) - q if (v % q) >= q // 2 else (v % q)) for v in out]
    return np.array(out, dtype=object if log_q is None else np.int64)


# ---------------------------------------------------------------------------
# rotations, sample extraction
# ---------------------------------------------------------------------------


def coef_rotation(poly, alpha: int) -> np.ndarray:
    """poly * X^alpha in Z[X]/(X^N+1) (schemas/polynomial_jax.py:150-169)."""
    poly = np.asarray(poly)
    n = poly.shape[-1]
    alpha = int(alpha)
    rot = np.roll(poly, alpha, axis=-1)
    div = alpha // n
    mod = alpha % n
    rot = rot * (-1) ** (div % 2)
    mask = np.arange(n) < mod
    return np.where(mask, -rot, rot)


def rotate_ciphertext(ct, exponent: int):
    """schemas/text_jax.py:348-362: rotate both components by X^exponent."""
    return coef_rotation(ct[0], exponent), coef_rotation(ct[1], exponent)


def sample_extract(ct, alpha: int):
    """schemas/text_jax.py:191-208: LWE (mask, body) of coefficient alpha."""
    a, b = np.asarray(ct[0]), np.asarray(ct[1])
    N = a.shape[-1]
    j = np.arange(N)
    src = np.where(j <= alpha, alpha - j, alpha - j + N)
    sign = np.where(j <= alpha, 1, -1)
    return a[..., src] * sign, b[..., alpha: alpha + 1]


# ---------------------------------------------------------------------------
# index preparation (cuTFHE/multiply.py:54-92)
# ---------------------------------------------------------------------------


def rotation_indices(pk, log_q: int, N: int, collapse: int, n_lut: int) -> np.ndarray:
    """Group rotation exponents of cuTFHE/multiply.py:67-79.

    val[g,p] = sum_j pk[g*c+j] * bit_j(p)           (float64, exact)
    idx = int32(round_half_even(2N*val/(q*n_lut)) * n_lut) mod 2N   (floor mod)
    pk: (..., n) float64 or integer.  Returns (..., n//c, 2^c) int32.
    """
    pk = np.asarray(pk, dtype=np.float64)
    n = pk.shape[-1]
    n_iter = n // collapse
    kron = all_binary_vectors(collapse).astype(np.float64)
    val = pk.reshape(*pk.shape[:-1], n_iter, collapse) @ kron.T
    q = 2.0 ** log_q
    t = (2 * N) * val / (q * n_lut)
    idx = (np.rint(t) * n_lut).astype(np.int32)
    return np.mod(idx, 2 * N).astype(np.int32)


def body_rotation(b, log_q: int, N: int, n_lut: int) -> np.ndarray:
    """b_ms = round(2N*b/(q*n_lut))*n_lut (schemas/RLWE_jax.py:295)."""
    q = 2.0 ** log_q
    b = np.asarray(b, dtype=np.float64)
    return (np.rint((2 * N) * b / (q * n_lut)) * n_lut).astype(np.int64)


def reshape_bsk_for_kernel(bsk_ordered, n_iter: int, collapse: int):
    """cuTFHE/multiply.py:81-86: (2,2,l,n_eq,N/2) -> (n_iter,2,l,2,2^c,N/2)."""
    shape = list(bsk_ordered.shape)
    shape[3] = n_iter
    shape.insert(4, 2 ** collapse)
    return np.transpose(np.reshape(bsk_ordered, shape), (3, 0, 2, 1, 4, 5))


# ---------------------------------------------------------------------------
# blind rotation: FP64 restatement of the reference kernel
# ---------------------------------------------------------------------------


def blind_rotate_f64(ct, bsk_k, idx, log_beta: int, log_q: int, L: int) -> np.ndarray:
    """Collapsed blind rotation exactly as csrc/mul_bootstrap.cuh:20-136 does it
    (also schemas/text_jax.py:390-435), batched over the leading axis.

    ct:    (B, 2, N) int64 (row 0 = mask A, row 1 = body B)
    bsk_k: (n_iter, 2[a], L, 2[L/R], m, N/2) complex128 (kernel layout)
    idx:   (B, n_iter, m) int32
    Per step: integer decomposition (dec_fft.cuh:4-31), forward transform of
    the 2L digit polys, acc_bc = sum_m BSK*ROT (mul_bootstrap.cuh:51-59),
    prod = sum_bc acc_bc*DEC (:61-66), inverse, x - q*rint(x/q), round to
    nearest even (:84-92).
    """
    ct = np.array(ct, dtype=np.int64)
    B, _, N = ct.shape
    n_iter = bsk_k.shape[0]
    m = bsk_k.shape[4]
    M = N // 2
    q = 2.0 ** log_q
    jj = np.arange(M)
    for g in range(n_iter):
        dec = decompose_int(ct, log_beta, log_q, L)              # (L, B, 2, N)
        dec = np.transpose(dec, (1, 0, 2, 3)).reshape(B, 2 * L, N)  # bc = 2*lev + comp
        dec_fft = negacyclic_fft(dec.astype(np.float64))           # (B, 2L, M)
        rot = zeta_pow(idx[:, g, :, None].astype(np.int64) * (2 * jj[None, None, :] + 1), N)  # (B,m,M)
        chunk = bsk_k[g].reshape(2, 2 * L, m, M)                   # (a, bc, m, d)
        acc = np.einsum('abmd,Bmd->Babd', chunk, rot)
        prod = np.einsum('Babd,Bbd->Bad', acc, dec_fft)
        val = negacyclic_ifft_raw(prod)
        val = val - q * np.rint(val / q)
        ct = np.rint(val).astype(np.int64)
    return ct


# ---------------------------------------------------------------------------
# keys and encryption (own seeded keys; the reference draws its keys with
# np.random.choice + jax threefry, which cannot be reproduced without JAX)
# ---------------------------------------------------------------------------


def uniform_torus(rng, shape, log_q: int) -> np.ndarray:
    q = 1 << log_q
    return rng.integers(-q // 2, q // 2, size=shape, dtype=np.int64)


def gaussian_noise(rng, shape, sigma: float) -> np.ndarray:
    return np.rint(rng.standard_normal(size=shape) * sigma).astype(np.int64)


def negacyclic_matrix(s) -> np.ndarray:
    """T with (a*s)_k = sum_j a_j T[j,k] in Z[X]/(X^N+1):
    T[j,k] = s[k-j] if k >= j else -s[k-j+N]."""
    s = np.asarray(s)
    N = len(s)
    j = np.arange(N)[:, None]
    k = np.arange(N)[None, :]
    d = k - j
    return np.where(d >= 0, s[np.mod(d, N)], -s[np.mod(d, N)])


def negacyclic_mul_binary(a, s) -> np.ndarray:
    """Exact a(X)*s(X) mod X^N+1 for |s_i| <= 1 and N*max|a_j| < 2^53,
    batched over leading axes of a.

    A float64 matrix product with the negacyclic matrix of s: every partial
    sum is an integer below N*max|a| < 2^53, so the result is exact whatever
    the summation order.
    """
    a = np.asarray(a, dtype=np.int64)
    N = a.shape[-1]
    assert float(np.abs(a).max(initial=0)) * N < 2.0 ** 53
    T = negacyclic_matrix(np.asarray(s, np.float64))
    out = a.reshape(-1, N).astype(np.float64) @ T
    return np.rint(out).astype(np.int64).reshape(a.shape)


def rlwe_encrypt(rng, m_poly, S, log_q: int, sigma: float):
    """schemas/RLWE_jax.py:36-54: b = A*S + m + e (mod q), A uniform."""
    m_poly = np.asarray(m_poly, dtype=np.int64)
    A = uniform_torus(rng, m_poly.shape, log_q)
    e = gaussian_noise(rng, m_poly.shape, sigma)
    b = centered_mod(negacyclic_mul_binary(A, S) + m_poly + e, log_q)
    return A, b


def rlwe_phase(A, b, S, log_q: int):
    return centered_mod(np.asarray(b, np.int64) - negacyclic_mul_binary(A, S), log_q)


def lwe_encrypt(rng, mu, s, log_q: int, sigma: float):
    """schemas/LWE_jax.py:9-28 (noise as round(N(0,1)*sigma))."""
    mu = np.asarray(mu, dtype=np.int64)
    n = len(s)
    a = uniform_torus(rng, mu.shape + (n,), log_q)
    e = gaussian_noise(rng, mu.shape, sigma)
    b = centered_mod(a @ np.asarray(s, np.int64) + mu + e, log_q)
    return a, b


def lwe_phase(a, b, s, log_q: int):
    """schemas/noise_estimators.py:7-11 without the message term."""
    return centered_mod(np.asarray(b, np.int64).reshape(-1) -
                        np.asarray(a, np.int64).reshape(len(np.asarray(b).reshape(-1)), -1)
                        @ np.asarray(s, np.int64), log_q)


def bsk_collapse_coef(rng, s_lwe, S, log_q: int, log_beta: int, L: int, collapse: int,
                      sigma: float):
    """Collapsed bootstrapping key in the coefficient domain.

    Follows schemas/RLWE_jax.py:162-197: for every group g and pattern p the
    indicator ind = 1[s_g == p]; the left RGSW rows encrypt -S*ind*q/beta^i,
    the right rows encrypt ind*q/beta^i on the constant coefficient.
    Returns int64 (2[L/R], 2[pk/b], n_eq, L, N): the array jnp.array(BSK)
    builds in nn_tfhe/activation.py:49.
    """
    s_lwe = np.asarray(s_lwe, np.int64)
    S = np.asarray(S, np.int64)
    N = len(S)
    kron = all_binary_vectors(collapse)
    groups = s_lwe.reshape(-1, collapse)
    all_eq = np.all(groups[:, None, :] == kron[None, :, :], axis=-1).astype(np.int64).ravel()
    n_eq = len(all_eq)
    out = np.empty((2, 2, n_eq, L, N), dtype=np.int64)
    for i in range(1, L + 1):
        g = 1 << (log_q - i * log_beta)          # q / beta^i
        left_msg = centered_mod(-S[None, :] * all_eq[:, None] * g, log_q)   # (n_eq, N)
        right_msg = np.zeros((n_eq, N), np.int64)
        right_msg[:, 0] = all_eq * g
        A, b = rlwe_encrypt(rng, left_msg, S, log_q, sigma)
        out[0, 0, :, i - 1], out[0, 1, :, i - 1] = A, b
        A, b = rlwe_encrypt(rng, right_msg, S, log_q, sigma)
        out[1, 0, :, i - 1], out[1, 1, :, i - 1] = A, b
    return out


def bsk_to_fourier_ordered(bsk_coef):
    """nn_tfhe/activation.py:49-53: FFT then permute [1,0,3,2,4] to
    (2[pk/b], 2[L/R], l, n_eq, N/2) complex128."""
    f = negacyclic_fft(bsk_coef.astype(np.float64))
    return np.transpose(f, (1, 0, 3, 2, 4))


def bsk_coef_kernel_layout(bsk_coef, n_iter: int, collapse: int):
    """Coefficient-domain BSK in the kernel layout (n_iter,2[a],l,2[L/R],m,N)."""
    ordered = np.transpose(bsk_coef, (1, 0, 3, 2, 4))      # (2,2,l,n_eq,N)
    return reshape_bsk_for_kernel(ordered, n_iter, collapse)


def keyswitch_key(rng, s_new, s_old, log_q: int, log_beta: int, l: int, sigma: float):
    """schemas/LWE_jax.py:86-93: KSK[j,i] = LWE_{s_new}(-s_old[i]*q/beta^(j+1)).

    Returns (ksk_a (l, N_in, n) int64, ksk_b (l, N_in) int64).
    """
    s_old = np.asarray(s_old, np.int64)
    msgs = np.stack([centered_mod(-s_old * (1 << (log_q - (j + 1) * log_beta)), log_q)
                     for j in range(l)])
    a, b = lwe_encrypt(rng, msgs, s_new, log_q, sigma)
    return a, b


def key_switch_lwe(ksk_a, ksk_b, a, b, log_q: int, log_beta: int, l: int):
    """schemas/LWE_jax.py:97-117 with gadget_product_LWE (:65-82).

    digits = float decomposition (polynomial_jax.py:90-111), exact here;
    a' = sum d*KSK_a, b' = b + sum d*KSK_b, centered mod q.  Batched over
    the leading axis of a (B, N_in), b (B,).
    """
    q = float(2 ** log_q)
    d = decompose_float(np.asarray(a, np.float64), float(2 ** log_beta), l, q)  # (l,B,N_in)
    d = np.transpose(d, (1, 0, 2)).astype(np.int64)                        # (B,l,N_in)
    B = d.shape[0]
    dflat = d.reshape(B, -1)
    a_new = centered_mod(dflat @ ksk_a.reshape(-1, ksk_a.shape[-1]), log_q)
    b_new = centered_mod(np.asarray(b, np.int64).reshape(-1) + centered_mod(dflat @ ksk_b.reshape(-1), log_q), log_q)
    return a_new, b_new


# ---------------------------------------------------------------------------
# PBS body (schemas/RLWE_jax.py:288-303) and LUT builders
# ---------------------------------------------------------------------------


def cuboot_merge_tail(br_out, b, log_q: int, n_lut: int):
    """Rotate the blind-rotated LUT by -b_ms and extract coefficients
    0..n_lut-1 (schemas/RLWE_jax.py:295-303).

    br_out (B,2,N) int64, b (B,) -> masks (B,n_lut,N), bodies (B,n_lut).
    """
    br_out = np.asarray(br_out, np.int64)
    B, _, N = br_out.shape
    bms = body_rotation(b, log_q, N, n_lut).reshape(-1)
    masks = np.empty((B, n_lut, N), np.int64)
    bodies = np.empty((B, n_lut), np.int64)
    for i in range(B):
        rot = rotate_ciphertext(br_out[i], -int(bms[i]))
        for k in range(n_lut):
            a_k, b_k = sample_extract(rot, k)
            masks[i, k] = a_k
            bodies[i, k] = b_k[0]
    return masks, bodies


def get_plain_LUT(degree_lut: int, n_lut: int, index: int, s: float, gamma: float):
    """nn_tfhe/utils_spike.py:72-87: sub-LUT `index` of a step 1[x>=gamma]
    sampled on linspace(-s, s) and interleaved with stride n_lut, then
    multiplied by X^(-N/2)."""
    sub = degree_lut // n_lut
    part = np.linspace(-s, s, num=sub, endpoint=False)
    vals = np.where(part >= gamma, 1.0, 0.0)
    full = np.zeros(degree_lut)
    full[np.arange(sub) * n_lut + index] = vals
    return coef_rotation(full, -degree_lut // 2)


def lut_from_function(N: int, n_lut: int, fns, msg_bits: int, delta_out: int):
    """Many-LUT test polynomial for signed messages x in [-2^(b-1), 2^(b-1))
    encoded as x*q/2^(b+1) (one padding bit).

    Sub-LUT k (k < n_lut) sits at coefficients base*n_lut + k; base index
    reads the signed phase psi = base*n_lut - N/2 (in units of q/2N), the
    message is round(psi / (N/2^b)); output delta_out*f_k(x).  The top half
    box (x = 2^(b-1)) holds -f(-2^(b-1)), the negacyclic image of the bottom
    one.  The layout is the one of get_plain_LUT (utils_spike.py:72-87),
    including the final X^(-N/2).
    """
    sub = N // n_lut
    W = N >> msg_bits
    full = np.zeros(N, np.int64)
    base = np.arange(sub)
    psi = base * n_lut - N // 2
    msg = np.floor(psi / W + 0.5).astype(np.int64)
    top = 1 << (msg_bits - 1)
    for k, f in enumerate(fns):
        vals = [(-f(int(v) - 2 * top) if v >= top else f(int(v))) for v in msg]
        full[base * n_lut + k] = delta_out * np.asarray(vals, np.int64)
    return coef_rotation(full, -N // 2)


def predicted_bs_std(N: int, n: int, log_beta: int, L: int, collapse: int,
                     sigma_bsk: float, log_q: int) -> float:
    """Reference's analytic blind-rotation noise std (noise_study.py:120),
    in torus units * q (beta_x factor left out)."""
    beta = 2.0 ** log_beta
    q = 2.0 ** log_q
    v = (2 * L * N * n * beta ** 2 * sigma_bsk ** 2 * 2 ** collapse / (collapse * 12)
         + N * q ** 2 * n / (collapse * beta ** (2 * L) * 24))
    return float(np.sqrt(v))
