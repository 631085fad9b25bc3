// pz_arith.hpp -- device integer arithmetic for the PBS path on gfx950.
//
// Every routine is the exact integer function the reference computes (same
// Shoup quotient, same lazy ranges, same decomposition carries), written for
// 64-wide wavefronts: 64-bit words live in VGPR pairs and the 64x64->128
// high product is emulated from 32-bit v_mad_u64_u32 / v_mul_hi_u32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;

#define PZ_DEV __device__ __forceinline__

// hi 64 bits of a*b (the Shoup quotient). __umul64hi lowers to three
// v_mad_u64_u32 plus one v_mul_hi_u32 on CDNA.
PZ_DEV u64 mulhi64(u64 a, u64 b) { return __umul64hi(a, b); }

// The two halves of the Shoup step, as 32-bit products for v_mad_u64_u32.
// shoup_quot: floor(ws*a / 2^64) (three middle/high partial products and the
// high half of the low one, carries kept exact).
PZ_DEV u64 shoup_quot(u64 a, u64 ws)
{
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), s0 = (u32)ws, s1 = (u32)(ws >> 32);
    const u64 t = (u64)s0 * a1 + __umulhi(s0, a0);
    const u64 u = (u64)s1 * a0 + (u32)t;
    return (u64)s1 * a1 + (t >> 32) + (u >> 32);
}
// c + w*a - q*h (mod 2^64), written as c + w*a + (2^64 - q)*h: the same word,
// with the high cross products folded into one 32-bit sum.  The four cross
// products form one chain of v_mad_u64_u32, each taking the previous result
// as its 64-bit addend (only the low word is used, so the addend's high word
// is free and no add merges them: 4 instructions instead of v_mul_lo_u32,
// v_mad, two v_mul_lo_u32 and v_add3), the products of w first, which do not
// wait for the quotient.  The empty asm keeps the compiler from narrowing the
// chain back to 32-bit multiplies and adds.
PZ_DEV u64 shoup_rem(u64 c, u64 a, u64 w, u64 h, u64 q)
{
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32);
    const u64 nq = 0 - q;
    const u32 h0 = (u32)h, h1 = (u32)(h >> 32);
    u64 mc = (u64)(u32)(w >> 32) * a0;
    mc = (u64)(u32)w * a1 + mc;
    mc = (u64)h1 * (u32)nq + mc;
    mc = (u64)h0 * (u32)(nq >> 32) + mc;
    asm("" : "+v"(mc));
    const u32 m = (u32)mc;
    const u64 lo = (u64)(u32)w * a0 + (u64)(u32)nq * h0 + c;
    // m << 32 only changes the high word: one 32-bit add (as a 64-bit add the
    // compiler spends a move on a zero low word as well)
    u32 hi = (u32)(lo >> 32);
    asm("v_add_u32 %0, %1, %2" : "=v"(hi) : "v"(hi), "v"(m));
    return ((u64)hi << 32) | (u32)lo;
}

// A quotient at most 2 below floor(ws*a / 2^64): the low partial product and
// the two carries out of the middle column are dropped (each fractional part
// is below 1).  c + w*a - q*h then lies in c + [0, 4q) instead of c + [0, 2q)
// and is the same residue; callers whose lazy ranges admit 4q use it.
PZ_DEV u64 shoup_quot4(u64 a, u64 ws)
{
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), s0 = (u32)ws, s1 = (u32)(ws >> 32);
    const u32 x = __umulhi(s0, a1), y = __umulhi(s1, a0);
    return (u64)s1 * a1 + ((u64)x + y);
}
// With the twin T = floor(w 2^63 / q) instead (T < 2^63, so its high word is
// below 2^31) and a < 2^63, the middle column a1*T0 + a0*T1 stays below 2^64
// and is summed exactly: Q = a1*T1 + floor((a1*T0 + a0*T1) / 2^32) drops only
// the low product and the column's low word (each below 1 after scaling) and
// T's truncation (a / 2^64 < 1/2), so floor(a*w / 2q) - 2 <= Q <= floor(a*w /
// 2q), and c + w*a - 2q*Q lies in c + [0, 6q): three v_mad_u64_u32 (and the
// move that zero-extends the column's high word) against shoup_quot4's two
// v_mul_hi_u32, a v_mad_u64_u32 and a 64-bit add.
PZ_DEV u64 shoup_quot6(u64 a, u64 T)
{
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), t0 = (u32)T, t1 = (u32)(T >> 32);
    const u64 col = (u64)a1 * t0 + (u64)a0 * t1;
    return (u64)a1 * t1 + (col >> 32);
}

// w*a mod q in [0, 4q) for any a < 2^64 (the two-step form with shoup_quot4)
PZ_DEV u64 shoup_mul4(u64 a, u64 w, u64 ws, u64 q)
{
    return shoup_rem(0, a, w, shoup_quot4(a, ws), q);
}

// Shoup multiply by 1 (ws = floor(2^64/q)), the fused inverse's reduction of
// an unreduced sum, for q > 2^32 so that ws < 2^32: the quotient needs two
// 32-bit products instead of the general four, the remainder v - q*h two.
// shoup_one: the exact quotient floor(v*ws/2^64) (v1*ws + hi(v0*ws), high
// word), the same word as shoup_mul(v, 1, ws, q), in [0, 2q).
// shoup_one4: the quotient shoup_quot4 forms when ws < 2^32 (hi(v1*ws)), the
// same word as shoup_mul4(v, 1, ws, q), in [0, 4q) (in fact [0, 3q)).
PZ_DEV u64 shoup_one_rem(u64 v, u32 h, u64 q)
{
    const u64 nq = 0 - q;  // v + (2^64 - q)*h mod 2^64
    const u64 lo = (u64)(u32)nq * h + v;
    u32 hi = (u32)(lo >> 32);
    const u32 m = (u32)(nq >> 32) * h;
    asm("v_add_u32 %0, %1, %2" : "=v"(hi) : "v"(hi), "v"(m));
    return ((u64)hi << 32) | (u32)lo;
}
PZ_DEV u64 shoup_one(u64 v, u32 ws, u64 q)
{
    const u64 t = (u64)(u32)(v >> 32) * ws + __umulhi((u32)v, ws);
    return shoup_one_rem(v, (u32)(t >> 32), q);
}
PZ_DEV u64 shoup_one4(u64 v, u32 ws, u64 q)
{
    return shoup_one_rem(v, __umulhi((u32)(v >> 32), ws), q);
}

// ShoupMul::mul, src/utils.rs:43-46: b*a - q*floor(b_shoup*a/2^64) (wrapping),
// result in [0, 2q) for any a < 2^64.
// SPLIT selects the two-step form above (two fewer VALU instructions per
// multiply, more registers live); both return the same word.
template <bool SPLIT = true>
PZ_DEV u64 shoup_mul(u64 a, u64 w, u64 ws, u64 q)
{
    if constexpr (SPLIT) return shoup_rem(0, a, w, shoup_quot(a, ws), q);
    else return w * a - q * mulhi64(ws, a);
}

// forward_butterly_0_to_4q, src/ntt.rs:31-49
template <bool SPLIT = true>
PZ_DEV void fb_4q(u64 &x, u64 &y, u64 w, u64 ws, u64 q, u64 q2)
{
    u64 xx = x >= q2 ? x - q2 : x;
    u64 t = shoup_mul<SPLIT>(y, w, ws, q);
    x = xx + t;
    y = xx + q2 - t;
}

PZ_DEV u64 umin(u64 a, u64 b) { return a < b ? a : b; }

// 2x + c (c uniform) as one v_lshl_add_u64: written as plain C the compiler
// turns (x << 1) + c into ((x + c/2) << 1), an add and a 64-bit shift
PZ_DEV u64 twice_plus(u64 x, u64 c)
{
    u64 r;
    asm("v_lshl_add_u64 %0, %1, 1, %2" : "=v"(r) : "v"(x), "s"(c));
    return r;
}

// forward_butterly_0_to_2q, src/ntt.rs:51-75
template <bool SPLIT = true>
PZ_DEV void fb_2q(u64 &x, u64 &y, u64 w, u64 ws, u64 q, u64 q2)
{
    u64 xx = x >= q2 ? x - q2 : x;
    u64 t = shoup_mul<SPLIT>(y, w, ws, q);
    u64 ox = xx + t, oy = xx - t;
    x = umin(ox, ox - q2);
    y = umin(oy, oy + q2);
}

// inverse_butterfly_0_to_2q, src/ntt.rs:84-104
template <bool SPLIT = true>
PZ_DEV void ib_2q(u64 &x, u64 &y, u64 w, u64 ws, u64 q, u64 q2)
{
    u64 xd = x + y;
    xd = xd >= q2 ? xd - q2 : xd;
    u64 t = x + q2 - y;
    x = xd;
    y = shoup_mul<SPLIT>(t, w, ws, q);
}

// ModularOpsU64 add_mod_fast / sub_mod_fast / neg (src/backend/modulus_u64.rs:46-77,
// :147-149)
PZ_DEV u64 add_mod(u64 a, u64 b, u64 q)
{
    u64 o = a + b;
    return o >= q ? o - q : o;
}
PZ_DEV u64 sub_mod(u64 a, u64 b, u64 q) { return a >= b ? a - b : (q + a) - b; }
PZ_DEV u64 neg_mod(u64 a, u64 q) { return q - a; }

// One Shoup multiply-accumulate step of shoup_matrix_fma
// (src/backend/modulus_u64.rs:252-260): out in [0,2q) -> [0,2q).
// REDUCE = false leaves out + [0,2q) unreduced (fused operators whose
// consumer accepts the wider range).
// QUOT4 takes the product with shoup_quot4 (out + [0, 4q), REDUCE = false only).
template <bool SPLIT = true, bool REDUCE = true, bool QUOT4 = false>
PZ_DEV u64 shoup_fma(u64 out, u64 a, u64 as, u64 b, u64 q, u64 q2)
{
    static_assert(!QUOT4 || (SPLIT && !REDUCE), "the [0,4q) product is left unreduced");
    u64 v;
    if constexpr (QUOT4) v = shoup_rem(out, b, a, shoup_quot4(b, as), q);
    else if constexpr (SPLIT) v = shoup_rem(out, b, a, shoup_quot(b, as), q);  // a*b + out - q*quot
    else v = a * b + out - q * mulhi64(as, b);
    if constexpr (!REDUCE) return v;
    return v >= q2 ? v - q2 : v;
}

// Barrett mul_mod_fast (src/backend/modulus_u64.rs:106-130) for keygen-side
// products. Constants: logq = 64 - clz(q+1), mu = 2^(2logq+3)/q, alpha = logq+3.
struct Barrett {
    u64 q;
    u32 logq, alpha;
    u64 mu_lo, mu_hi;
};

PZ_DEV void mul128(u64 a, u64 b, u64 &lo, u64 &hi)
{
    lo = a * b;
    hi = __umul64hi(a, b);
}

PZ_DEV u64 barrett_mul(const Barrett &br, u64 a, u64 b)
{
    u64 ab_lo, ab_hi;
    mul128(a, b, ab_lo, ab_hi);
    // tmp = ab >> (logq-2)
    u32 s = br.logq - 2;
    u64 t_lo = (ab_lo >> s) | (ab_hi << (64 - s));
    u64 t_hi = ab_hi >> s;
    // k = (tmp * mu) >> (alpha + 2); tmp*mu < 2^128 by the reference's bounds
    u64 p0_lo, p0_hi, p1_lo, p1_hi, p2_lo, p2_hi;
    mul128(t_lo, br.mu_lo, p0_lo, p0_hi);
    mul128(t_lo, br.mu_hi, p1_lo, p1_hi);
    mul128(t_hi, br.mu_lo, p2_lo, p2_hi);
    u64 mid = p0_hi + p1_lo;
    u64 c1 = mid < p0_hi;
    u64 mid2 = mid + p2_lo;
    u64 c2 = mid2 < mid;
    u64 top = p1_hi + p2_hi + t_hi * br.mu_hi + c1 + c2;
    // (p0_lo, mid2, top) is the product's low 192 bits; k = product >> (alpha+2)
    // (only its low 64 bits matter: the result is taken mod 2^64)
    u32 sh = br.alpha + 2;  // in (0, 64) for 2 < logq <= 60
    u64 k_lo = (p0_lo >> sh) | (mid2 << (64 - sh));
    (void)top;
    // out = (ab - k*q) mod 2^64
    u64 out = ab_lo - k_lo * br.q;
    return out >= br.q ? out - br.q : out;
}

// Gadget decomposer (src/decomposer.rs:158-178, :217-234, :264-309).
struct Decomp {
    u64 q;
    u32 logb, d, ignore_bits;
    // the d = 1 digit's addends (decomp_digit_d1): rc, rc - q, rc - 2q (mod
    // 2^64), rc = 2^(ignore_bits-1) (0 without ignored bits); set by
    // decomp_centring on the host
    u64 cadd0, cadd1, cadd2;
};

inline void decomp_centring(Decomp &dc)
{
    const u64 rc = dc.ignore_bits ? 1ull << (dc.ignore_bits - 1) : 0;
    dc.cadd0 = rc;
    dc.cadd1 = rc - dc.q;
    dc.cadd2 = rc - 2 * dc.q;
}

// Digit `r` (0-based) of the signed decomposition of v. Release-build
// semantics: for v == q the !(q - v) + 1 wraps to 0.
PZ_DEV u64 decomp_prepare(const Decomp &dc, u64 value)
{
    if (value >= (dc.q >> 1)) value = ~(dc.q - value) + 1;
    u32 ib = dc.ignore_bits;
    if (ib != 0) {
        u64 msb = (value & ((1ull << ib) - 1)) >> (ib - 1);
        value = (value >> ib) + msb;
    }
    return value;
}

// One DecomposerIter::next step; returns the digit and advances `value`.
PZ_DEV u64 decomp_next(const Decomp &dc, u64 &value)
{
    u64 bmask = (1ull << dc.logb) - 1;
    u64 bby2 = 1ull << (dc.logb - 1);
    u64 k = value & bmask;
    value = (value - k) >> dc.logb;
    u64 carry = (k > bby2) || (k == bby2 && (value & 1) == 1);
    u64 neg_carry = 0 - carry;
    value = value + carry;
    return (neg_carry & dc.q) + k - (carry << dc.logb);
}

// Digit 0 alone (what a d = 1 decomposition uses), for logb <= 30: only the
// low 32 bits of the rounded value matter, and the carry into the next digit
// is not formed.  Same word as decomp_digit(dc, v, 0).
PZ_DEV u64 decomp_digit0(const Decomp &dc, u64 v)
{
    if (v >= (dc.q >> 1)) v -= dc.q;  // -(q - v) in two's complement; v == q gives 0
    // round(v / 2^ib) = (floor(v / 2^(ib-1)) + 1) / 2, in 32 bits: bits 0..30
    // of the result are exact, and the digit and its carry use bits 0..logb
    const u32 ib = dc.ignore_bits;
    const u32 sh = ib ? ib - 1 : 0, rnd = ib ? 1u : 0u;  // uniform
    const u32 value = ((u32)(v >> sh) + rnd) >> rnd;
    const u32 b = 1u << dc.logb, half = b >> 1;
    const u32 k = value & (b - 1);
    // carry when k > B/2, or k == B/2 and the next digit is odd: one compare
    // of k + (that bit) against B/2, no divergent branch
    const bool carry = k + ((value >> dc.logb) & 1u) > half;
    return (u64)k + (carry ? dc.q - b : 0);  // q + k - B for a negative digit
}

// Digit of a d = 1 decomposition as q + digit, a representative in
// [q - B/2, q + B/2] of the reference's digit (the consumer is a transform
// that accepts inputs below 2q).  With d = 1 the decomposer ignores
// bitlen(q) - logb bits, so the rounded signed value lies in [-B/2, B/2] and
// the carry rule never changes it: the digit is the value itself
// (negative values as q + value; the value B/2 with an even next digit stays
// B/2, the value -B/2 carries into q - B/2).  Needs logb + ignore_bits <= 63
// and ignore_bits > 0.  ca0, ca1, ca2: the decomposer's cadd0, cadd1, cadd2
// (2^(ignore_bits - 1) minus 0, q and 2q), passed as values: a select between
// fields of a struct is turned into a load from a stack copy at a selected
// offset.
// LOOSE takes v in [0, 2q] (a representative the reference would hold as
// v mod q): v_s = v - 2q from q + floor(q/2) on, the reference's rule on
// v - q there.
template <bool LOOSE = false>
PZ_DEV u64 decomp_digit_d1(u64 q, u32 ignore_bits, u64 ca0, u64 ca1, u64 ca2, u64 v)
{
    // v_s + 2^(ib-1) in two's complement, v_s = v - q when v >= floor(q/2):
    // the addend is selected among the three precomputed ones (computed here
    // from q, the compiler selects the multiple of q, subtracts it from v and
    // adds 2^(ib-1) after it: a 64-bit subtraction more per digit)
    const u64 h = q >> 1;
    u64 c = v >= h ? ca1 : ca0;
    if constexpr (LOOSE) c = v >= q + h ? ca2 : c;
    const u64 t = v + c;
    return q + (u64)((int64_t)t >> ignore_bits);  // floor, then + q
}

PZ_DEV u64 decomp_digit(const Decomp &dc, u64 v, u32 r)
{
    u64 value = decomp_prepare(dc, v);
    u64 dg = 0;
    for (u32 i = 0; i <= r; i++) dg = decomp_next(dc, value);
    return dg;
}
