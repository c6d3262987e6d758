// Integer torus helpers shared by the gfx950 kernels: classes mod
// q = 2^lq (lq <= 64) as int64 centred representatives, sums wrapping mod
// 2^64, and the signed digit rule of the reference's float decomposition
// restated exactly on integers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bslstm {

// centred representative mod 2^lq of a 64-bit class
__device__ __forceinline__ int64_t tmod(uint64_t x, int lq)
{
    const int s = 64 - lq;
    return (int64_t)(x << s) >> s;
}

// round_half_even(r / 2^s) for s in [1, 63], without overflow
__device__ __forceinline__ int64_t rhe_shift(int64_t r, int s)
{
    const int64_t fl = r >> s;
    const uint64_t rem = (uint64_t)r & ((1ULL << s) - 1);
    const uint64_t half = 1ULL << (s - 1);
    return fl + (int64_t)((rem > half) | ((rem == half) & ((uint64_t)fl & 1)));
}

// Next signed digit of the reference's float decomposition
// (polynomial_jax.py:90-111: x = gamma/q; d = round(x*beta); x = x*beta - d)
// restated exactly on the centred integer residual r: d = round_half_even(
// r / 2^s), r -= d*2^s, with s = log2_q - level*log2_beta (levels from 1).
__device__ __forceinline__ int64_t next_digit(int64_t &r, int s)
{
    if (s <= 0) {
        const int64_t d = r;
        r = 0;
        return d;
    }
    const int64_t d = rhe_shift(r, s);
    r = (int64_t)((uint64_t)r - ((uint64_t)d << s));
    return d;
}

// Residual of that recursion before the level whose digit has weight
// 2^(S - log2_beta), straight from the centred class r (lq bits): with
// S = log2_q - level*log2_beta it is r's centred remainder mod 2^S, the tie
// rem = 2^(S-1) going negative when bit S of r is set (the earlier digits'
// round-half-even never changes the bits below their own weight, and bit S
// is the parity the last rounding looked at).  S >= lq: r itself; S <= 0:
// 0 (the recursion has emptied it).  Lets a thread start the recursion at
// any level.
__device__ __forceinline__ int64_t residual_at(int64_t r, int S, int lq)
{
    if (S >= lq)
        return r;
    if (S <= 0)
        return 0;
    const uint64_t rem = (uint64_t)r & ((1ULL << S) - 1);
    const uint64_t half = 1ULL << (S - 1);
    const bool up = rem > half || (rem == half && (((uint64_t)r >> S) & 1));
    return up ? (int64_t)(rem - (1ULL << S)) : (int64_t)rem;
}

// The same digits without the recursion, read from bit fields of the centred
// class u (lq bits, sign-extended).  up(S) is the round-half-even carry of the
// residual at weight 2^S -- 1 iff residual_at(r, S) is the negative
// representative: at S = lq the sign of the class; S <= 0: 0; else bit S-1
// set and (some lower bit set, or bit S set: the tie goes to even).  With
// S = lq - j*lb and S1 = max(S - lb, 0), digit j of the recursion is
//   d_j = bits [S1, S) of u + up(S1) - up(S) * 2^(S - S1)
// (the residuals before and after level j differ by d_j * 2^S1).
//
// carries_of(u) holds every carry at once: bit p is up(p + 1) for p < lq.
// "Some bit below p" is "p above the lowest set bit of u", so
//   bit p = u_p & (u_(p+1) | p > ctz(u)),
// and bit lq-1 is the sign, which the formula gives through the sign-extended
// bit lq except at lq = 64 (u = 2^63): or-ed in.  One word per coefficient
// per level loop; each digit is then a field and two carry bits.
__device__ __forceinline__ uint64_t carries_of(uint64_t u, int lq)
{
    const uint64_t low = u & (0 - u);              // lowest set bit (0 for u = 0)
    const uint64_t above = ~(low | (low - 1));     // positions above it
    return (u & ((u >> 1) | above)) | (u & (1ULL << (lq - 1)));
}

// Digit j for a run of coefficients: the level's shifts and masks are the
// same for every coefficient (wave-uniform), so they are worked out once
// (digit_level) and each digit (digit_at_level) is two 64-bit shifts and four
// 32-bit operations (v_bfi, v_bfe_i32, and, add), with no branch.  With
// w = S - S1, uu = u >> S1 (the field in bits [0, w)) and cc = cv >> (S1 - 1)
// (up(S1) in bit 0, up(S) in bit w), inserting bits [0, w) of uu into cc and
// reading bits [0, w] as a signed number gives field - up(S) * 2^w; bit 0 of
// cc adds up(S1).  At S1 = 0, up(S1) = 0 and cc = cv << 1.  Needs w <= 30
// (log2 of the base).
struct DigitLevel {
    int kind;           // 0: every digit is 0 (j >= l or S <= 0); 1: S1 > 0; 2: S1 = 0
    int s1, w;
    uint32_t fmask;     // bits [0, w)
};

__device__ __forceinline__ DigitLevel digit_level(int j, int l, int lb, int lq)
{
    const int S = lq - j * lb;
    if (j >= l || S <= 0)
        return {0, 0, 0, 0u};
    const int S1 = S > lb ? S - lb : 0;
    const int w = S - S1;
    return {S1 > 0 ? 1 : 2, S1, w, (1u << w) - 1u};
}

template <int KIND>
__device__ __forceinline__ int digit_at_level(const DigitLevel &d, uint64_t u, uint64_t cv)
{
    static_assert(KIND == 1 || KIND == 2, "kind 0 digits are 0");
    uint32_t uu, cc;
    if constexpr (KIND == 1) {
        uu = (uint32_t)(u >> d.s1);
        cc = (uint32_t)(cv >> (d.s1 - 1));
    } else {
        uu = (uint32_t)u;
        cc = (uint32_t)cv << 1;
    }
    const uint32_t x = (uu & d.fmask) | (cc & ~d.fmask);
    return __builtin_amdgcn_sbfe((int)x, 0, d.w + 1) + (int)(cc & 1u);
}

// Torus class of a rounded FP64 transform output (jax_inversefourier's
// round, then centred_mod q).
__device__ __forceinline__ int64_t f64_to_torus(double v, double q, double inv_q, int lq)
{
    v = rint(v - q * rint(v * inv_q));
    if (lq >= 64)
        return v >= 0x1p63 ? INT64_MIN : (int64_t)v;
    return tmod((uint64_t)(int64_t)v, lq);
}

}  // namespace bslstm
