// pz_ntt.hpp -- block-level negacyclic NTT over a 54-bit prime for gfx950.
//
// One ring element of N = 2^LOGN u64 coefficients is held by N/8 threads,
// 8 coefficients per thread in VGPR pairs.  The 11 (for N=2048) radix-2
// stages of the reference are grouped into radix-8 phases of three stages
// that run entirely in registers; between phases the polynomial is
// exchanged through LDS (padded one word in nine to spread banks).
//
// Forward (Cooley-Tukey, src/ntt.rs:110-180) takes coefficients in the
// "stride" layout (thread j holds x = j + (N/8)*r) and leaves the evaluation
// vector in the "contiguous" layout (thread j holds x = 8j + r).  The inverse
// (Gentleman-Sande, src/ntt.rs:187-278) runs the phases backwards.  Each
// element sees exactly the reference's butterfly sequence and twiddles, so
// every lazy intermediate value is identical to the reference.
#pragma once
#include "pz_arith.hpp"

namespace pz {

__host__ __device__ constexpr int ilog_max1(int v) { return v > 0 ? v : 1; }

// LDS index of coefficient x (one pad word every 8)
PZ_DEV u32 lds_ix(u32 x) { return x + (x >> 3); }

template <int LOGN>
struct Geom {
    static constexpr int N = 1 << LOGN;
    static constexpr int TPB = N / 8;
    static constexpr int LDS_WORDS = N + N / 8;  // per polynomial
};

// Stride of the thread's R coefficients (R = 8: radix-8 phases of three
// stages; R = 4: radix-4 phases of two) in the phase starting at stage S0.
template <int R>
struct RadixLog {
    static constexpr int value = R == 8 ? 3 : 2;
};
template <int LOGN, int S0, int R = 8>
struct Phase {
    static constexpr int N = 1 << LOGN;
    static constexpr int LR = RadixLog<R>::value;
    static constexpr int ST = ilog_max1(N >> (S0 + LR));
    static constexpr int BS = R * ST;
    static constexpr int NS = (LOGN - S0) < LR ? (LOGN - S0) : LR;
    PZ_DEV static u32 base(u32 j) { return (j / ST) * BS + (j % ST); }
    PZ_DEV static u32 x(u32 j, int r) { return base(j) + ST * r; }
};

// Load through an address_space(1) pointer: table pointers that went through
// an asm barrier are generic, and generic (flat) loads also count against the
// LDS counter, serialising every twiddle load with the LDS exchanges.
typedef u64 u64x2 __attribute__((ext_vector_type(2)));

PZ_DEV u64x2 gld2(const u64x2 *p)
{
    return *(const __attribute__((address_space(1))) u64x2 *)(p);
}

// NTT tables in device memory: (w, w_shoup) pairs interleaved so one 16-byte
// load fetches both, powers in the reference's bit-reversed order.
struct Tables {
    const u64x2 *__restrict__ fw;  // forward  (psi^bo, shoup)
    const u64x2 *__restrict__ fw63;  // forward  (psi^bo, floor(psi^bo 2^63 / q)): FWD_FAST6
    const u64x2 *__restrict__ iw63;  // inverse  (psi^-bo, floor(psi^-bo 2^63 / q)): INV6 stages
    const u64x2 *__restrict__ iw;  // inverse  (psi^-bo, shoup)
    u64 n_inv, n_inv_s;
    u64 q, q2;
    u64 one_s;  // floor(2^64 / q): Shoup twin of 1, reduces any word to [0, 2q)
};

// Forward modes: the reference's forward_lazy ([0,2q)) and forward ([0,q)),
// and FWD_FAST for the fused blind rotation: no conditional subtractions, so
// stage k leaves values below (2k+1)q (< 23q < 2^59 for N = 2048).  FAST is
// only used where a Shoup product consumes the result (Shoup accepts any
// 64-bit input), so every ciphertext stays identical to the reference.
// FWD_FAST4 is FWD_FAST with the [0,4q) product (shoup_mul4) and y' = x + 4q - t:
// from inputs below 2q (the split step's digits, q + digit) stage k leaves
// values below (4k+2)q (< 46q < 2^60 for N = 2048).
// FWD_FAST6 is FWD_FAST4 with the [0,6q) product of shoup_quot6 (twins from
// tb.fw63, modulus 2q) and y' = x + 6q - t: from inputs below 2q stage k
// leaves values below (6k+2)q (< 68q < 2^61 for N = 2048), inside the
// quotient's a < 2^63.
enum { FWD_LAZY = 0, FWD_STRICT = 1, FWD_FAST = 2, FWD_FAST4 = 3, FWD_FAST6 = 4 };

// Unreduced tail of the IB = 8 fast inverse (see inv_stage).
#ifndef PZ_INV_TAIL4
#define PZ_INV_TAIL4 1
#endif
// The IB = 8 fast inverse's products in the [0,6q) form (shoup_quot6, twins
// from tb.iw63, modulus 2q; see inv_stage) in every stage but KR = 7 and the
// last
#ifndef PZ_INV6
#define PZ_INV6 1
#endif
// executed stage k = LOGN - S (1-based) of an IB = 8 fast inverse takes the
// 2^63 twin: before KR its difference is below 2^(k+3) q <= 2^9 q < 2^63 and
// [0,6q) is within the next stage's bound 2^k 8q; in the unreduced tail (KR <
// k < LOGN) inputs are below 2^(k-6) q, the difference below 2^(k-5) q, and the
// next stage's bound 2^(k-5) q >= 8q holds [0,6q); stage KR's difference can
// reach 2^10 q, beyond 2^63
template <int LOGN, int S>
constexpr bool inv6_stage() { return PZ_INV6 && (LOGN - S < 7 || (LOGN - S > 7 && S > 0)); }

// One forward stage S inside a phase of stride ST. KIND: 0 = 0_to_4q
// butterfly, 1 = last-stage 0_to_2q (lazy), 2 = last stage + reduce to [0,q),
// 3 = unreduced (FWD_FAST), 4 = unreduced with the [0,4q) product (FWD_FAST4),
// 5 = unreduced with the [0,6q) product (FWD_FAST6).
// SM selects the two-step Shoup form
// (shoup_mul<true>, fewer instructions); the one-step form has fewer live
// registers, which keeps the one-work-group blind rotation within 256 VGPRs.
template <int LOGN, int S, int ST, int NP, int KIND, bool SM, int R, bool PRE = false>
PZ_DEV void fwd_stage(u64 (&v)[NP][R], u32 base, const Tables &tb, const u64x2 *pre = nullptr)
{
    constexpr int N = 1 << LOGN;
    constexpr int T = N >> (S + 1);
    constexpr int LD = T / ST;
    constexpr int M = 1 << S;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r & LD) continue;
        const u32 x = base + ST * r;
        const u32 i = x / (2 * T);
        const u64x2 tw = PRE ? pre[r] : gld2((KIND == 5 ? tb.fw63 : tb.fw) + M + i);
        const u64 w = tw.x, ws = tw.y;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if (KIND == 0) {
                fb_4q<SM>(v[p][r], v[p][r + LD], w, ws, tb.q, tb.q2);
            } else if (KIND == 3) {
                const u64 x = v[p][r];
                const u64 t = shoup_mul<SM>(v[p][r + LD], w, ws, tb.q);
                v[p][r] = x + t;
                v[p][r + LD] = x + tb.q2 - t;
            } else if (KIND == 4) {
                // x + t with x as the remainder's addend (no separate add), and
                // x + 4q - t as 2x + 4q - (x + t)
                const u64 x = v[p][r], a = v[p][r + LD];
                const u64 xt = shoup_rem(x, a, w, shoup_quot4(a, ws), tb.q);
                v[p][r] = xt;
                v[p][r + LD] = twice_plus(x, tb.q2 << 1) - xt;
            } else if (KIND == 5) {
                // the same with t in [0, 6q): the remainder modulo 2q, and
                // x + 6q - t as 2x + 6q - (x + t)
                const u64 x = v[p][r], a = v[p][r + LD];
                const u64 xt = shoup_rem(x, a, w, shoup_quot6(a, ws), tb.q2);
                v[p][r] = xt;
                v[p][r + LD] = twice_plus(x, tb.q2 * 3) - xt;
            } else {
                fb_2q<SM>(v[p][r], v[p][r + LD], w, ws, tb.q, tb.q2);
                if (KIND == 2) {
                    v[p][r] = umin(v[p][r], v[p][r] - tb.q);
                    v[p][r + LD] = umin(v[p][r + LD], v[p][r + LD] - tb.q);
                }
            }
        }
    }
}

// One inverse stage S (m = 2^S, t = N >> (S+1)); S == 0 is the last stage,
// which also scales by n^-1 when SCALE (and reduces to [0,q) when STRICT).
// SCALE = false is for products whose key rows already carry the n^-1
// factor: the result is the same residue, and with STRICT the same word.
//
// FAST (fused operators, inputs below IB*q with IB = 2, 4 or 8, no n^-1):
// executed stage k = LOGN - S receives values below 2^(k-1) IB q.  Early
// stages skip the reduction of the sum, x' = x + y, and take
// y' = Shoup(x + 2^(k-1) IB q - y); stage KR (9, 8, 7 for IB = 2, 4, 8)
// reduces x' = x + y < 2^KR IB q = 2^10 q to [0,2q) with a Shoup multiply by
// 1 (2^10 q < 2^64 for q < 2^54); later stages are the reference butterfly
// again.  With IB = 8 the products of stages before KR come from shoup_mul4
// ([0,4q), within the next stage's bound 2^k IB q), and the tail is unreduced
// as well (PZ_INV_TAIL4): stage KR leaves [0,4q) (shoup_mul4 by 1 for the
// sum), stage k > KR receives values below 2^(k-KR+1) q and adds that bound
// in its difference, and only the last stage reduces, its sum with an exact
// Shoup multiply by 1 and both outputs from [0,2q) to [0,q).  Without the
// tail stage KR's products are exact and the reference butterflies follow.
// Residues are the reference's throughout, and the strict last stage returns
// the same words.
template <int LOGN, int S, int ST, int NP, bool STRICT, bool SCALE, bool SM, bool FAST, int IB,
          int R, bool PRE = false>
PZ_DEV void inv_stage(u64 (&v)[NP][R], u32 base, const Tables &tb, const u64x2 *pre = nullptr)
{
    constexpr int N = 1 << LOGN;
    constexpr int T = N >> (S + 1);
    constexpr int LD = T / ST;
    constexpr int M = 1 << S;
    constexpr int K = LOGN - S;  // executed stage, 1-based
    constexpr int LIB = IB == 8 ? 3 : (IB == 4 ? 2 : 1);  // log2 IB
    constexpr int KR = 10 - LIB;
    constexpr bool T4 = PZ_INV_TAIL4 && FAST && IB == 8;  // unreduced tail
    // N = 2048 contexts have q > 2^32 (pz_ctx_create), so floor(2^64/q) fits
    // 32 bits and the reductions by 1 take the short forms
    constexpr bool ONE32 = LOGN >= 11;
    static_assert(!FAST || !SCALE, "the fast inverse leaves n^-1 to the key rows");
    static_assert(IB == 2 || ((IB == 4 || IB == 8) && FAST),
                  "inputs below 4q or 8q need the fast inverse");
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r & LD) continue;
        const u32 x = base + ST * r;
        const u32 i = x / (2 * T);
        constexpr bool I6 = FAST && IB == 8 && inv6_stage<LOGN, S>();
        const u64x2 tw = PRE ? pre[r] : gld2((I6 ? tb.iw63 : tb.iw) + M + i);
        const u64 w = tw.x, ws = tw.y;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if constexpr (FAST && K <= KR) {
                const u64 xv = v[p][r], yv = v[p][r + LD];
                const u64 sum = xv + yv;
                const u64 dif = xv + (tb.q << (K - 1 + LIB)) - yv;
                if constexpr (T4 && K == KR) {
                    v[p][r] = ONE32 ? shoup_one4(sum, (u32)tb.one_s, tb.q)
                                    : shoup_mul4(sum, 1, tb.one_s, tb.q);
                    v[p][r + LD] = shoup_mul4(dif, w, ws, tb.q);
                } else {
                    v[p][r] = K != KR ? sum
                                      : (ONE32 ? shoup_one(sum, (u32)tb.one_s, tb.q)
                                               : shoup_mul<SM>(sum, 1, tb.one_s, tb.q));
                    if constexpr (I6)  // dif < 2^(K+3) q <= 2^9 q < 2^63; [0,6q) <= 2^K IB q
                        v[p][r + LD] = shoup_rem(0, dif, w, shoup_quot6(dif, ws), tb.q2);
                    else
                        v[p][r + LD] = (IB == 8 && K < KR) ? shoup_mul4(dif, w, ws, tb.q)
                                                           : shoup_mul<SM>(dif, w, ws, tb.q);
                }
            } else if constexpr (T4) {
                const u64 xv = v[p][r], yv = v[p][r + LD];
                const u64 sum = xv + yv;
                const u64 dif = xv + (tb.q << (K - KR + 1)) - yv;  // inputs below 2^(K-KR+1) q
                if constexpr (S == 0) {  // the last stage: exact, [0,2q) (-> [0,q) if STRICT)
                    const u64 xo = ONE32 ? shoup_one(sum, (u32)tb.one_s, tb.q)
                                         : shoup_mul<SM>(sum, 1, tb.one_s, tb.q);
                    const u64 yo = shoup_mul<SM>(dif, w, ws, tb.q);
                    v[p][r] = STRICT ? umin(xo, xo - tb.q) : xo;
                    v[p][r + LD] = STRICT ? umin(yo, yo - tb.q) : yo;
                } else {
                    v[p][r] = sum;
                    if constexpr (I6)  // [0,6q) within the next stage's 2^(K-KR+2) q >= 8q
                        v[p][r + LD] = shoup_rem(0, dif, w, shoup_quot6(dif, ws), tb.q2);
                    else
                        v[p][r + LD] = shoup_mul4(dif, w, ws, tb.q);
                }
            } else {
                ib_2q<SM>(v[p][r], v[p][r + LD], w, ws, tb.q, tb.q2);
            }
            if (S == 0 && !SCALE && STRICT && !T4) {
                v[p][r] = umin(v[p][r], v[p][r] - tb.q);
                v[p][r + LD] = umin(v[p][r + LD], v[p][r + LD] - tb.q);
            } else if (S == 0 && SCALE) {
                u64 ox = shoup_mul<SM>(v[p][r], tb.n_inv, tb.n_inv_s, tb.q);
                u64 oy = shoup_mul<SM>(v[p][r + LD], tb.n_inv, tb.n_inv_s, tb.q);
                if (STRICT) {
                    ox = umin(ox, ox - tb.q);
                    oy = umin(oy, oy - tb.q);
                }
                v[p][r] = ox;
                v[p][r + LD] = oy;
            }
        }
    }
}

template <bool F63 = false>
PZ_DEV struct Tables fresh(const Tables &tb);

// The twiddles of one phase, [stage in the phase][r] (only entries with
// !(r & LD) are used; the loads of equal indices are one load).  Loaded ahead
// of the LDS exchange that precedes the phase (PF transforms): the exchange's
// fence keeps the compiler from moving a load above it, and loaded after it
// each stage's twiddle latency is exposed once the register budget stops the
// scheduler from issuing several together.
template <int R>
struct PhaseTw {
    u64x2 t[3][R];
};
// TAB: 0 = forward (tb.fw), 1 = inverse (tb.iw), 2 = forward with 2^63
// twins (tb.fw63, FWD_FAST6), 3 = inverse of an IB = 8 fast transform (tb.iw63
// for its stages before KR, tb.iw after)
template <int LOGN, int S0, int R, int TAB>
PZ_DEV void phase_tw(PhaseTw<R> &pt, u32 j, const Tables &tb)
{
    using P = Phase<LOGN, S0, R>;
    const u32 base = P::base(j);
#pragma unroll
    for (int e = 0; e < P::NS; e++) {
        const int S = S0 + e;
        const u64x2 *tab = TAB == 2                                  ? tb.fw63
                         : TAB == 3 && PZ_INV6 && (LOGN - S < 7 || (LOGN - S > 7 && S > 0)) ? tb.iw63
                         : TAB == 1 || TAB == 3                      ? tb.iw
                                                                     : tb.fw;
        const int T = (1 << LOGN) >> (S + 1);
        const int LD = T / P::ST;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r & LD) continue;
            pt.t[e][r] = gld2(tab + (1 << S) + (base + P::ST * r) / (2 * T));
        }
    }
}

// Forward phase starting at stage S0 (up to log2 R stages); PRE: twiddles
// from pt.
template <int LOGN, int S0, int NP, int MODE, bool SM, int R, bool PRE = false>
PZ_DEV void fwd_phase(u64 (&v)[NP][R], u32 j, const Tables &tb, const PhaseTw<R> *pt = nullptr)
{
    using P = Phase<LOGN, S0, R>;
    const u32 base = P::base(j);
    constexpr int LAST = LOGN - 1;
    constexpr int K_LAST = MODE == FWD_STRICT ? 2
                         : MODE == FWD_LAZY   ? 1
                         : MODE == FWD_FAST   ? 3
                         : MODE == FWD_FAST4  ? 4
                                              : 5;
    constexpr int K_MID = MODE == FWD_FAST ? 3 : MODE == FWD_FAST4 ? 4 : MODE == FWD_FAST6 ? 5 : 0;
    fwd_stage<LOGN, S0, P::ST, NP, (S0 == LAST) ? K_LAST : K_MID, SM, R, PRE>(
        v, base, tb, PRE ? pt->t[0] : nullptr);
    if constexpr (P::NS > 1)
        fwd_stage<LOGN, S0 + 1, P::ST, NP, (S0 + 1 == LAST) ? K_LAST : K_MID, SM, R, PRE>(
            v, base, tb, PRE ? pt->t[1] : nullptr);
    if constexpr (P::NS > 2)
        fwd_stage<LOGN, S0 + 2, P::ST, NP, (S0 + 2 == LAST) ? K_LAST : K_MID, SM, R, PRE>(
            v, base, tb, PRE ? pt->t[2] : nullptr);
}

// Inverse phase covering stages S0+NS-1 .. S0 (descending).
template <int LOGN, int S0, int NP, bool STRICT, bool SCALE, bool SM, bool FAST, int IB, int R,
          bool PRE = false>
PZ_DEV void inv_phase(u64 (&v)[NP][R], u32 j, const Tables &tb, const PhaseTw<R> *pt = nullptr)
{
    using P = Phase<LOGN, S0, R>;
    const u32 base = P::base(j);
    if constexpr (P::NS > 2)
        inv_stage<LOGN, S0 + 2, P::ST, NP, STRICT, SCALE, SM, FAST, IB, R, PRE>(
            v, base, tb, PRE ? pt->t[2] : nullptr);
    if constexpr (P::NS > 1)
        inv_stage<LOGN, S0 + 1, P::ST, NP, STRICT, SCALE, SM, FAST, IB, R, PRE>(
            v, base, tb, PRE ? pt->t[1] : nullptr);
    inv_stage<LOGN, S0, P::ST, NP, STRICT, SCALE, SM, FAST, IB, R, PRE>(
        v, base, tb, PRE ? pt->t[0] : nullptr);
}

// Store the thread's coefficients (layout of phase SA) into LDS, barrier,
// load them back in the layout of phase SB, barrier.
// TAIL = false leaves out the second barrier.  That is safe where the next
// LDS write of every thread goes only to the words the thread itself read
// here (the next exchange of the same transform stores the layout SB that was
// just loaded, and a transform that follows stores its first layout where the
// last one was loaded); any other writer must synchronise first.
//
// Wave-local exchanges.  In a phase of stride ST = N >> (S + log2 R) <= 64,
// wave w of a polynomial's N/R threads holds exactly the words
// [64 R w, 64 R (w + 1)).  An exchange between two such phases stays within
// each wave, and LDS operations of one wave execute in order, so with
// TAIL = false it needs no work-group barrier: a wave barrier keeps the
// compiler from moving the loads above the stores.
template <int LOGN, int SA, int SB, int R = 8>
constexpr bool wave_local_exchange()
{
    return (SA < SB ? SA : SB) >= LOGN - 6 - RadixLog<R>::value;
}

PZ_DEV void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LOGN, int SA, int SB, int NP, int R, bool TAIL>
PZ_DEV void exchange(u64 (&v)[NP][R], u32 j, u64 *lds)
{
    constexpr int W = Geom<LOGN>::LDS_WORDS;
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int r = 0; r < R; r++) lds[p * W + lds_ix(Phase<LOGN, SA, R>::x(j, r))] = v[p][r];
    if constexpr (!TAIL && wave_local_exchange<LOGN, SA, SB, R>()) wave_sync();
    else __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int r = 0; r < R; r++) v[p][r] = lds[p * W + lds_ix(Phase<LOGN, SB, R>::x(j, r))];
    if constexpr (TAIL) __syncthreads();
}

// Twiddle-table pointers passed through an empty asm before each phase.  The
// compiler then cannot hoist the (loop-invariant) twiddle loads out of the
// blind-rotation loop: hoisted, they pin ~160 VGPRs per thread and cap the
// kernel at one wave per SIMD.
#ifndef PZ_TWIDDLE_RELOAD
#define PZ_TWIDDLE_RELOAD 1
#endif
// The pointers are uniform; readfirstlane states it where the compiler cannot
// prove it, so the asm's scalar-register operands always hold.
PZ_DEV const u64x2 *uniform_ptr(const u64x2 *p)
{
    const u64 v = (u64)p;
    const u32 lo = (u32)__builtin_amdgcn_readfirstlane((int)(u32)v);
    const u32 hi = (u32)__builtin_amdgcn_readfirstlane((int)(u32)(v >> 32));
    return (const u64x2 *)(((u64)hi << 32) | lo);
}
// F63: the FWD_FAST6 table as well (passed through the asm only where it is
// read: one more pointer there cost the transform kernels 40 VGPRs)
template <bool F63>
PZ_DEV Tables fresh(const Tables &tb)
{
    Tables t = tb;
#if PZ_TWIDDLE_RELOAD
    t.fw = uniform_ptr(t.fw);
    t.iw = uniform_ptr(t.iw);
    asm volatile("" : "+s"(t.fw), "+s"(t.iw));
    if constexpr (F63) {
        t.fw63 = uniform_ptr(t.fw63);
        t.iw63 = uniform_ptr(t.iw63);
        asm volatile("" : "+s"(t.fw63), "+s"(t.iw63));
    }
#endif
    return t;
}

// Forward NTT of NP polynomials: stride layout in, contiguous layout out.
// MODE FWD_STRICT = Ntt::forward (output in [0,q)), FWD_LAZY = forward_lazy
// ([0,2q)), FWD_FAST = unreduced (see above).
// Phases S0, S0 + log2 R, ... with an LDS exchange between consecutive ones.
// TAIL = false: exchanges without their second barrier (see exchange).
// PF: each phase's twiddles come from a PhaseTw loaded before the exchange
// that precedes the phase (the first phase's from the caller's pt).
template <int LOGN, int S0, int NP, int MODE, bool SM, int R, bool TAIL, bool PF = false>
PZ_DEV void fwd_from(u64 (&v)[NP][R], u32 j, u64 *lds, const Tables &tb,
                     const PhaseTw<R> *pt = nullptr)
{
    constexpr int LR = RadixLog<R>::value;
    if constexpr (PF) fwd_phase<LOGN, S0, NP, MODE, SM, R, true>(v, j, tb, pt);
    else fwd_phase<LOGN, S0, NP, MODE, SM, R>(v, j, fresh<MODE == FWD_FAST6>(tb));
    if constexpr (S0 + LR < LOGN) {
        if constexpr (PF) {
            PhaseTw<R> nt;
            phase_tw<LOGN, S0 + LR, R, MODE == FWD_FAST6 ? 2 : 0>(nt, j,
                                                                  fresh<MODE == FWD_FAST6>(tb));
            __builtin_amdgcn_sched_barrier(0);  // issued ahead of the exchange
            exchange<LOGN, S0, S0 + LR, NP, R, TAIL>(v, j, lds);
            fwd_from<LOGN, S0 + LR, NP, MODE, SM, R, TAIL, PF>(v, j, lds, tb, &nt);
        } else {
            exchange<LOGN, S0, S0 + LR, NP, R, TAIL>(v, j, lds);
            fwd_from<LOGN, S0 + LR, NP, MODE, SM, R, TAIL>(v, j, lds, tb);
        }
    }
}

// Work-group barriers of one forward transform with TAIL = false: its
// exchanges that are not wave-local.
template <int LOGN, int R = 8>
constexpr int ntt_fwd_barriers()
{
    constexpr int LR = RadixLog<R>::value;
    int n = 0;
    for (int s0 = 0; s0 + LR < LOGN; s0 += LR)
        if (s0 < LOGN - 6 - LR) n++;
    return n;
}

template <int LOGN, int NP, int MODE, bool SM = true, int R = 8, bool TAIL = true>
PZ_DEV void ntt_fwd(u64 (&v)[NP][R], u32 j, u64 *lds, const Tables &tb)
{
    static_assert(LOGN >= 9 && LOGN <= 12, "ring size");
    static_assert((MODE != FWD_FAST && MODE != FWD_FAST4 && MODE != FWD_FAST6) || LOGN <= 11,
                  "unreduced values must stay below 2^64 (FWD_FAST6: below 2^63)");
    static_assert(R == 8 || R == 4, "8 or 4 coefficients per thread");
    fwd_from<LOGN, 0, NP, MODE, SM, R, TAIL>(v, j, lds, tb);
}

// The same with every phase's twiddles loaded ahead of its exchange; pt holds
// the first phase's (fwd_tw_first), which the caller issues early.
template <int LOGN, int R, int MODE = FWD_FAST4>
PZ_DEV void fwd_tw_first(PhaseTw<R> &pt, u32 j, const Tables &tb)
{
    phase_tw<LOGN, 0, R, MODE == FWD_FAST6 ? 2 : 0>(pt, j, fresh<MODE == FWD_FAST6>(tb));
}
template <int LOGN, int NP, int MODE, bool SM = true, int R = 8, bool TAIL = true>
PZ_DEV void ntt_fwd_pf(u64 (&v)[NP][R], u32 j, u64 *lds, const Tables &tb, const PhaseTw<R> &pt)
{
    static_assert(LOGN >= 9 && LOGN <= 12, "ring size");
    static_assert((MODE != FWD_FAST && MODE != FWD_FAST4 && MODE != FWD_FAST6) || LOGN <= 11,
                  "unreduced values must stay below 2^64 (FWD_FAST6: below 2^63)");
    static_assert(R == 8 || R == 4, "8 or 4 coefficients per thread");
    fwd_from<LOGN, 0, NP, MODE, SM, R, TAIL, true>(v, j, lds, tb, &pt);
}

// Inverse NTT of NP polynomials: contiguous layout in, stride layout out.
// STRICT = Ntt::backward ([0,q)), else backward_lazy ([0,2q)); SCALE = false
// leaves out the n^-1 factor; FAST = unreduced early stages (see inv_stage).
// Phases S0, S0 - log2 R, ..., 0 with an LDS exchange between consecutive ones.
template <int LOGN, int S0, int NP, bool STRICT, bool SCALE, bool SM, bool FAST, int IB, int R,
          bool TAIL, bool PF = false>
PZ_DEV void inv_from(u64 (&v)[NP][R], u32 j, u64 *lds, const Tables &tb,
                     const PhaseTw<R> *pt = nullptr)
{
    constexpr int LR = RadixLog<R>::value;
    if constexpr (PF) inv_phase<LOGN, S0, NP, STRICT, SCALE, SM, FAST, IB, R, true>(v, j, tb, pt);
    else
        inv_phase<LOGN, S0, NP, STRICT, SCALE, SM, FAST, IB, R>(
            v, j, fresh<FAST && IB == 8 && PZ_INV6>(tb));
    if constexpr (S0 > 0) {
        if constexpr (PF) {
            PhaseTw<R> nt;
            phase_tw<LOGN, S0 - LR, R, FAST && IB == 8 ? 3 : 1>(
                nt, j, fresh<FAST && IB == 8 && PZ_INV6>(tb));
            __builtin_amdgcn_sched_barrier(0);  // issued ahead of the exchange
            exchange<LOGN, S0, S0 - LR, NP, R, TAIL>(v, j, lds);
            inv_from<LOGN, S0 - LR, NP, STRICT, SCALE, SM, FAST, IB, R, TAIL, PF>(v, j, lds, tb,
                                                                                 &nt);
        } else {
            exchange<LOGN, S0, S0 - LR, NP, R, TAIL>(v, j, lds);
            inv_from<LOGN, S0 - LR, NP, STRICT, SCALE, SM, FAST, IB, R, TAIL>(v, j, lds, tb);
        }
    }
}

template <int LOGN, int NP, bool STRICT, bool SCALE = true, bool SM = true, bool FAST = false,
          int IB = 2, int R = 8, bool TAIL = true>
PZ_DEV void ntt_inv(u64 (&v)[NP][R], u32 j, u64 *lds, const Tables &tb)
{
    static_assert(!FAST || ((STRICT || (IB == 8 && PZ_INV_TAIL4)) && LOGN <= 12),
                  "fast inverse: strict output, or [0,2q) from the unreduced tail");
    static_assert(R == 8 || R == 4, "8 or 4 coefficients per thread");
    constexpr int LR = RadixLog<R>::value;
    inv_from<LOGN, ((LOGN - 1) / LR) * LR, NP, STRICT, SCALE, SM, FAST, IB, R, TAIL>(v, j, lds,
                                                                                    tb);
}

// The inverse with every phase's twiddles loaded ahead of its exchange; pt
// holds the first phase's (inv_tw_first).
// I6: for an IB = 8 fast inverse (the 2^63 twins of its early stages)
template <int LOGN, int R, bool I6 = false>
PZ_DEV void inv_tw_first(PhaseTw<R> &pt, u32 j, const Tables &tb)
{
    constexpr int LR = RadixLog<R>::value;
    phase_tw<LOGN, ((LOGN - 1) / LR) * LR, R, I6 ? 3 : 1>(pt, j, fresh<I6 && PZ_INV6>(tb));
}
template <int LOGN, int NP, bool STRICT, bool SCALE = true, bool SM = true, bool FAST = false,
          int IB = 2, int R = 8, bool TAIL = true>
PZ_DEV void ntt_inv_pf(u64 (&v)[NP][R], u32 j, u64 *lds, const Tables &tb, const PhaseTw<R> &pt)
{
    static_assert(!FAST || ((STRICT || (IB == 8 && PZ_INV_TAIL4)) && LOGN <= 12),
                  "fast inverse: strict output, or [0,2q) from the unreduced tail");
    static_assert(R == 8 || R == 4, "8 or 4 coefficients per thread");
    constexpr int LR = RadixLog<R>::value;
    inv_from<LOGN, ((LOGN - 1) / LR) * LR, NP, STRICT, SCALE, SM, FAST, IB, R, TAIL, true>(
        v, j, lds, tb, &pt);
}

// Coefficient index of register r in the stride (coefficient-domain) layout
// and in the contiguous (evaluation-domain) layout.
template <int LOGN, int R = 8>
PZ_DEV u32 coef_x(u32 j, int r) { return j + (u32)((1 << LOGN) / R) * r; }
template <int LOGN, int R = 8>
PZ_DEV u32 eval_x(u32 j, int r) { return R * j + r; }

}  // namespace pz
