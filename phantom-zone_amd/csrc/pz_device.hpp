// pz_device.hpp -- device-side context and the two RLWE operators of the
// blind rotation as work-group-level routines (one RLWE per work group of
// N/8 threads, coefficients resident in VGPRs).
#pragma once
#include "pz_ntt.hpp"

namespace pz {

constexpr int MAX_AUTO = 32;

// Blind-rotation op list entries: an LWE index (RLWE x RGSW), or SCHED_AUTO
// plus an automorphism key index.  Length bound: n products plus at most one
// automorphism per bucket (2 (br_q/4 - 1) + 1 < br_q/2).
constexpr u32 SCHED_AUTO = 0x8000;
__host__ __device__ inline u32 br_sched_len(u32 br_q, u32 n) { return n + br_q / 2 + 4; }

// Forward NTT mode of the fused RLWE operators: FWD_FAST (unreduced) or
// FWD_LAZY (the reference's forward_lazy values).  Both give identical
// ciphertexts; only the evaluation-domain scratch values differ.
#ifndef PZ_FWD_MODE
#define PZ_FWD_MODE FWD_FAST
#endif

// Unreduced early inverse stages in the fused operators (inputs in [0,2q),
// strict outputs; requires q < 2^54).
#ifndef PZ_INV_FAST
#define PZ_INV_FAST 1
#endif
// Products of the split blind rotation left unreduced (accumulators below
// 4q; the fast inverse then takes inputs below 4q).
#ifndef PZ_LAZY_MAC
#define PZ_LAZY_MAC 1
#endif
#ifndef PZ_SPLIT_SM
#define PZ_SPLIT_SM true
#endif

// Split blind rotation with the [0,4q) Shoup product (shoup_quot4) in the
// forward transform, the key products and the inverse's unreduced stages:
// accumulators below 8q, FWD_FAST4 and the IB = 8 fast inverse.
#ifndef PZ_QUOT4
#define PZ_QUOT4 1
#endif
// The split step's digit transform with the [0,6q) products of FWD_FAST6 (one
// VALU instruction fewer per butterfly than FWD_FAST4; its outputs, below 68q,
// feed only the key products, which take any 64-bit input)
#ifndef PZ_FWD6
#define PZ_FWD6 1
#endif

#ifndef PZ_MAC_FENCE
#define PZ_MAC_FENCE 0
#endif

// Split step transforms with each phase's twiddles loaded ahead of the LDS
// exchange before the phase (the first phases' before the digits and before
// the second key product).
// Bits: 1 forward; 2 inverse, its first phase's twiddles issued before the
// second key product; 4 inverse, its first phase's issued at its start.
// PZ_TW_PREFETCH for 8 coefficients per thread, PZ_TW_PREFETCH4 for 4.
// Measured at B = 4096 (8 coefficients): forward 1.9 % faster, with the
// inverse (4) 4.1 %; with 2 the inverse twiddles and the second key row are
// live together and the step spills (21 % slower).  One bootstrap
// (4 coefficients): 8 % faster with 5.
#ifndef PZ_TW_PREFETCH
#define PZ_TW_PREFETCH 5
#endif
#ifndef PZ_TW_PREFETCH4
#define PZ_TW_PREFETCH4 5
#endif

// Split-step digits are q + digit (decomp_digit_d1, 9 VALU instructions) instead
// of the reference's representative in [0, q) (decomp_digit0, 15): the same
// residue, for the unreduced transform.
// The blind rotation's RLWE stays in [0, 2q) between steps.
// Wave priorities of the split step in the throughput variant (0: none)
#ifndef PZ_STEP_PRIO
#define PZ_STEP_PRIO 1
#endif
#ifndef PZ_LOOSE_STEPS
#define PZ_LOOSE_STEPS 1
#endif

// Everything a kernel needs about the parameter set (passed by value).
struct DevCtx {
    Tables tb;
    Decomp dec_a, dec_b, dec_auto, dec_lwe;
    Decomp dec_uitos;  // u_j -> s RLWE key switch (non-interactive sets; d = 0: none)
    u32 da, db, dauto, lwe_d;
    u32 n_lwe, br_q, w, g, embedding;
    u32 lwe_q_log;
    u64 lwe_mask;
    double lwe_qf, rlwe_qf;
    u64 qby4;
    Barrett br;
    const u32 *dlog;        // [br_q]
    const u64 *test_vec;    // [6][br_q/2]
    u64 auto_el[MAX_AUTO];  // automorphism exponents, reduced into [0, 2N)
};

// Device key (reference row layout, evaluation domain, Shoup twins apart).
// RGSW and automorphism rows are stored times n^-1 mod q, so the inverse NTT
// of a key product needs no final scaling.
struct DevKey {
    const u64 *rgsw, *rgsw_s;  // [n][rgsw_rows][N]
    const u64 *auto_k, *auto_s;  // [w+1][auto_rows][N]
    const void *ksk;           // [N*d][ldk], u16 or u32, columns n+1..ldk-1 zero
    // the KSK as KS_PLANES 7-bit planes for the MFMA key switch,
    // [plane][ldk][N*d] int8 (column-major: one column's rows contiguous), or null
    const int8_t *ksk_planes;
    u32 rgsw_rows, auto_rows, ksk_bytes, ldk;
};
// 7-bit planes of a KSK word (q_ks <= 2^21)
constexpr int KS_PLANES = 3;

// A DevKey read from memory is the same for the whole work group: keep its
// fields in scalar registers.
PZ_DEV u32 ufl(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }
template <class T>
PZ_DEV T *ufl_ptr(T *p)
{
    const u64 v = (u64)p;
    return (T *)(((u64)ufl((u32)(v >> 32)) << 32) | ufl((u32)v));
}
PZ_DEV DevKey uniform_key(const DevKey &k)
{
    DevKey u;
    u.rgsw = ufl_ptr(k.rgsw);
    u.rgsw_s = ufl_ptr(k.rgsw_s);
    u.auto_k = ufl_ptr(k.auto_k);
    u.auto_s = ufl_ptr(k.auto_s);
    u.ksk = ufl_ptr(k.ksk);
    u.ksk_planes = ufl_ptr(k.ksk_planes);
    u.rgsw_rows = ufl(k.rgsw_rows);
    u.auto_rows = ufl(k.auto_rows);
    u.ksk_bytes = ufl(k.ksk_bytes);
    u.ldk = ufl(k.ldk);
    return u;
}

// 8 consecutive words (evaluation-domain slice of one thread)
PZ_DEV void load8(const u64 *__restrict__ p, u64 (&v)[8])
{
    const ulonglong2 *p2 = reinterpret_cast<const ulonglong2 *>(p);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        ulonglong2 t = p2[i];
        v[2 * i] = t.x;
        v[2 * i + 1] = t.y;
    }
}

// out[r] = shoup_fma(out[r], K[8j+r], Ks[8j+r], D[r]) for the thread's slice
template <bool SPLIT>
PZ_DEV void fma_row(u64 (&acc)[8], const u64 *__restrict__ K, const u64 *__restrict__ Ks,
                    const u64 (&D)[8], u32 j, u64 q, u64 q2)
{
#if PZ_MAC_FENCE
    // keep this row's key loads from being hoisted above the previous row /
    // the NTT: hoisted, all rows' loads are live at once (~128 VGPRs)
    __builtin_amdgcn_sched_barrier(0);
#endif
    u64 kv[8], ks[8];
    load8(K + 8 * j, kv);
    load8(Ks + 8 * j, ks);
#pragma unroll
    for (int r = 0; r < 8; r++) acc[r] = shoup_fma<SPLIT>(acc[r], kv[r], ks[r], D[r], q, q2);
}

// acc[r] += K[8j+r] * D[r] as a [0,4q) Shoup product, unreduced: after n
// rows from zero the accumulator is below 4nq (< 2^62 for n <= 64 rows and
// q < 2^54); fma_done brings it to [0,2q) with a Shoup multiply by 1.
PZ_DEV void fma_row_lazy4(u64 (&acc)[8], const u64 *__restrict__ K, const u64 *__restrict__ Ks,
                          const u64 (&D)[8], u32 j, u64 q, u64 q2)
{
    u64 kv[8], ks[8];
    load8(K + 8 * j, kv);
    load8(Ks + 8 * j, ks);
#pragma unroll
    for (int r = 0; r < 8; r++) acc[r] = shoup_fma<true, false, true>(acc[r], kv[r], ks[r], D[r], q, q2);
}
PZ_DEV u64 fma_done(u64 a, const Tables &tb) { return shoup_mul<true>(a, 1, tb.one_s, tb.q); }

// The thread's slice of one key row and its Shoup twins (R consecutive words
// each), as 16-byte loads.
template <int R>
struct KeySlice {
    u64x2 k[R / 2], s[R / 2];
};
template <int R>
PZ_DEV void load_slice(KeySlice<R> &ks, const u64 *__restrict__ K, const u64 *__restrict__ Ks,
                       u32 j)
{
    // global (not flat) loads: flat loads also wait on the LDS counter
    typedef const __attribute__((address_space(1))) u64x2 *gp;
    const gp k2 = (gp)(K + R * j), s2 = (gp)(Ks + R * j);
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
        ks.k[i] = k2[i];
        ks.s[i] = s2[i];
    }
}
// acc[r] += D[r] * K[r] (unreduced, [0,4q) products: shoup_fma<.., QUOT4>)
template <bool SPLIT, bool REDUCE, bool QUOT4, int R>
PZ_DEV void fma_slice(u64 (&acc)[R], const KeySlice<R> &ks, const u64 (&D)[R], u64 q, u64 q2)
{
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
        acc[2 * i] = shoup_fma<SPLIT, REDUCE, QUOT4>(acc[2 * i], ks.k[i].x, ks.s[i].x, D[2 * i], q, q2);
        acc[2 * i + 1] =
            shoup_fma<SPLIT, REDUCE, QUOT4>(acc[2 * i + 1], ks.k[i].y, ks.s[i].y, D[2 * i + 1], q, q2);
    }
}

// rlwe_by_rgsw_shoup (src/rgsw/runtime.rs:754-855) on the RLWE held in
// registers (stride layout). K/Ks point at the RGSW's first row; rows are
// [A(-sm) x da, B(-sm) x da, A(m) x db, B(m) x db], N words each.
template <int LOGN>
PZ_DEV void ext_product_blk(u64 (&a)[8], u64 (&b)[8], const u64 *__restrict__ K,
                            const u64 *__restrict__ Ks, bool trivial, const DevCtx &c,
                            u32 j, u64 *lds)
{
    constexpr int N = 1 << LOGN;
    const u64 q = c.tb.q, q2 = c.tb.q2;
    u64 acc[2][8];
#pragma unroll
    for (int r = 0; r < 8; r++) acc[0][r] = acc[1][r] = 0;
    const u32 da = c.da, db = c.db;
    if (da == 1 && db == 1 && !trivial) {
        u64 D[2][8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            D[0][r] = decomp_digit0(c.dec_a, a[r]);
            D[1][r] = decomp_digit0(c.dec_b, b[r]);
        }
        ntt_fwd<LOGN, 2, PZ_FWD_MODE, false>(D, j, lds, c.tb);
        fma_row<false>(acc[0], K + 0 * N, Ks + 0 * N, D[0], j, q, q2);
        fma_row<false>(acc[0], K + 2 * N, Ks + 2 * N, D[1], j, q, q2);
        fma_row<false>(acc[1], K + 1 * N, Ks + 1 * N, D[0], j, q, q2);
        fma_row<false>(acc[1], K + 3 * N, Ks + 3 * N, D[1], j, q, q2);
    } else {
        if (!trivial) {
            for (u32 rr = 0; rr < da; rr++) {
                u64 D[1][8];
#pragma unroll
                for (int r = 0; r < 8; r++) D[0][r] = decomp_digit(c.dec_a, a[r], rr);
                ntt_fwd<LOGN, 1, PZ_FWD_MODE, false>(D, j, lds, c.tb);
                fma_row<false>(acc[0], K + (size_t)rr * N, Ks + (size_t)rr * N, D[0], j, q, q2);
                fma_row<false>(acc[1], K + (size_t)(da + rr) * N, Ks + (size_t)(da + rr) * N, D[0], j,
                        q, q2);
            }
        }
        for (u32 rr = 0; rr < db; rr++) {
            u64 D[1][8];
#pragma unroll
            for (int r = 0; r < 8; r++) D[0][r] = decomp_digit(c.dec_b, b[r], rr);
            ntt_fwd<LOGN, 1, PZ_FWD_MODE, false>(D, j, lds, c.tb);
            size_t o0 = (size_t)(2 * da + rr) * N, o1 = (size_t)(2 * da + db + rr) * N;
            fma_row<false>(acc[0], K + o0, Ks + o0, D[0], j, q, q2);
            fma_row<false>(acc[1], K + o1, Ks + o1, D[0], j, q, q2);
        }
    }
    ntt_inv<LOGN, 2, true, false, false, PZ_INV_FAST>(acc, j, lds, c.tb);  // n^-1: key rows
#pragma unroll
    for (int r = 0; r < 8; r++) {
        a[r] = acc[0][r];
        b[r] = acc[1][r];
    }
}

// rlwe_auto_shoup (src/rgsw/runtime.rs:530-643) with the automorphism map of
// generate_auto_map (src/rgsw/keygen.rs:16-41) computed on the fly:
// to = i*k mod 2N, negated when it wraps past N.
template <int LOGN>
PZ_DEV void auto_blk(u64 (&a)[8], u64 (&b)[8], const u64 *__restrict__ K,
                     const u64 *__restrict__ Ks, u64 kexp, bool trivial, const DevCtx &c,
                     u32 j, u64 *lds)
{
    constexpr int N = 1 << LOGN;
    constexpr int W = Geom<LOGN>::LDS_WORDS;
    const u64 q = c.tb.q, q2 = c.tb.q2;
    const u64 SIGN = 1ull << 63;
    if (trivial) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            u32 x = coef_x<LOGN>(j, r);
            u32 to = (u32)(((u64)x * kexp) & (2 * N - 1));
            bool pos = to < (u32)N;
            to = pos ? to : to - N;
            lds[lds_ix(to)] = pos ? b[r] : neg_mod(b[r], q);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; r++) b[r] = lds[lds_ix(coef_x<LOGN>(j, r))];
        __syncthreads();
        return;
    }
#pragma unroll
    for (int r = 0; r < 8; r++) {
        u32 x = coef_x<LOGN>(j, r);
        u32 to = (u32)(((u64)x * kexp) & (2 * N - 1));
        bool pos = to < (u32)N;
        to = pos ? to : to - N;
        lds[lds_ix(to)] = pos ? a[r] : neg_mod(a[r], q);
        lds[W + lds_ix(to)] = pos ? b[r] : (b[r] | SIGN);
    }
    __syncthreads();
    u64 bp[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        a[r] = lds[lds_ix(coef_x<LOGN>(j, r))];
        bp[r] = lds[W + lds_ix(coef_x<LOGN>(j, r))];
    }
    __syncthreads();
    u64 acc[2][8];
#pragma unroll
    for (int r = 0; r < 8; r++) acc[0][r] = acc[1][r] = 0;
    const u32 d = c.dauto;
    for (u32 rr = 0; rr < d; rr++) {
        u64 D[1][8];
#pragma unroll
        for (int r = 0; r < 8; r++) D[0][r] = decomp_digit(c.dec_auto, a[r], rr);
        ntt_fwd<LOGN, 1, PZ_FWD_MODE, false>(D, j, lds, c.tb);
        fma_row<false>(acc[0], K + (size_t)rr * N, Ks + (size_t)rr * N, D[0], j, q, q2);
        fma_row<false>(acc[1], K + (size_t)(d + rr) * N, Ks + (size_t)(d + rr) * N, D[0], j, q, q2);
    }
    ntt_inv<LOGN, 2, true, false, false, PZ_INV_FAST>(acc, j, lds, c.tb);  // n^-1: key rows
#pragma unroll
    for (int r = 0; r < 8; r++) {
        u64 v = bp[r] & ~SIGN;
        bool neg = (bp[r] & SIGN) != 0;
        a[r] = acc[0][r];
        b[r] = neg ? sub_mod(acc[1][r], v, q) : add_mod(acc[1][r], v, q);
    }
}

// ---------------------------------------------------------------------------
// Split work group (2 x N/R threads, R = 8 or 4 coefficients per thread) for
// d_a = d_b = d_auto = 1 (every NI parameter set): half h (wave-uniform)
// holds polynomial h of the RLWE (0 = a, 1 = b), R coefficients per thread in
// the stride layout, and exchanges through its own LDS region (lds + h*W).
// Half the registers of the one-half operators above, twice the waves per
// bootstrap (R = 4: a quarter of the registers, four times the waves, radix-4
// register phases).
// ---------------------------------------------------------------------------

// The work-group barriers of one ntt_fwd without exchange tails, for a half
// that has no transform to run.
template <int LOGN, int R>
PZ_DEV void ntt_fwd_idle()
{
#pragma unroll
    for (int e = 0; e < ntt_fwd_barriers<LOGN, R>(); e++) __syncthreads();
}

// One blind-rotation step on a split work group, either
//  - rlwe_by_rgsw_shoup (src/rgsw/runtime.rs:754-855) with d_a = d_b = 1:
//    each half decomposes and transforms its own polynomial, the halves swap
//    their digit transforms through LDS, half 0 accumulates a' over rows
//    A(-sm), A(m) and half 1 accumulates b' over rows B(-sm), B(m); or
//  - rlwe_auto_shoup (src/rgsw/runtime.rs:530-643) with d_auto = 1: each half
//    permutes its polynomial (half 1 into `stash`, W words of LDS nothing
//    else touches, where the permuted b waits for the final add); half 0
//    decomposes and transforms auto(a) while half 1 waits at the matching
//    barriers; half 0 accumulates a' (row A), half 1 takes the digit
//    transform, accumulates b' (row B) and adds the permuted b.  Trivial:
//    only b is permuted (q - v where the map negates, the reference's neg).
// Both steps share one instruction stream (the transforms and the row
// products appear once), which keeps the step within 128 VGPRs.  A trivial
// RLWE needs no special case in the product: its digit transform is 0 mod q.
// The accumulation order of the product differs from the reference's; the
// strict inverse returns the same words.
// The transforms run without exchange tails: each thread's first LDS store of
// a transform goes to the words it loaded last (the inverse ends in the
// stride layout the forward starts from, the forward ends in the contiguous
// layout the digit store and the inverse start from), so only the scattered
// stores of an automorphism's permutation need a barrier before them.  The
// other half reads this half's digit transform, so a barrier separates that
// read from the inverse's first store.
//
// TRIV = false (the blind rotation) runs a trivial RLWE's automorphism as a
// general one: its a part is zero, so the key products are zero and the result
// has the same residues, while zero coefficients that the map negates come out
// as 0 instead of the reference's q (q - 0).  The next RLWE x RGSW step reduces
// both alike, and a blind rotation always ends in non-trivial steps, so every
// output word is the same; the loop then has one path.  The batched operator
// returns the RLWE itself and keeps the exact form (TRIV = true).
template <int LOGN, int R, bool TRIV = true>
PZ_DEV void step_split(u64 (&p)[R], const u64 *__restrict__ K, const u64 *__restrict__ Ks,
                       bool is_auto, u64 kexp, bool trivial_in, const DevCtx &c, u32 j, u32 h,
                       u64 *lds, u64 *stash)
{
    const bool trivial = TRIV && trivial_in;
    constexpr int N = 1 << LOGN;
    constexpr int W = Geom<LOGN>::LDS_WORDS;
    const u64 q = c.tb.q, q2 = c.tb.q2;
    // the blind rotation keeps its RLWE in [0, 2q) between steps (the
    // inverse's last reduction to [0, q) left out); the batched operator
    // returns the reference's words
    constexpr bool LOOSE = !TRIV && PZ_LOOSE_STEPS;
    const u64 qn = LOOSE ? q2 : q;  // negation: qn - v
    const u64 SIGN = 1ull << 63;
    u64 *mine = lds + h * W;
    const u64 *theirs = lds + (h ^ 1) * W;
    if (is_auto) {
        // the previous transform's last loads from this region are done
        __syncthreads();
        u64 *perm = h ? stash : mine;
        const bool move = !(trivial && h == 0);
        // a wraps with a negation (and so does b when trivial); a non-trivial b
        // only records the sign, applied when it is added back.  The choice is
        // uniform, so each loop body is branch-free.
        if (move && (h == 0 || trivial)) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                u32 x = coef_x<LOGN, R>(j, r);
                u32 to = (u32)(((u64)x * kexp) & (2 * N - 1));
                const bool pos = to < (u32)N;
                to = pos ? to : to - N;
                const u64 v = p[r];
                perm[lds_ix(to)] = pos ? v : qn - v;
            }
        } else if (move) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                u32 x = coef_x<LOGN, R>(j, r);
                u32 to = (u32)(((u64)x * kexp) & (2 * N - 1));
                const bool pos = to < (u32)N;
                to = pos ? to : to - N;
                const u64 v = p[r];
                perm[lds_ix(to)] = pos ? v : (v | SIGN);
            }
        }
        __syncthreads();
        if (trivial) {  // uniform over the work group
            if (move) {
#pragma unroll
                for (int r = 0; r < R; r++) p[r] = perm[lds_ix(coef_x<LOGN, R>(j, r))];
            }
            __syncthreads();
            return;
        }
        if (h == 0) {
#pragma unroll
            for (int r = 0; r < R; r++) p[r] = mine[lds_ix(coef_x<LOGN, R>(j, r))];
        }
        __syncthreads();  // read before the transform rewrites the region
    }
    // Wave priorities with two work groups per CU (8 coefficients per thread),
    // which arbitrate between the two bootstraps on a CU: an automorphism step
    // (the shorter step, half of whose waves idle through its transform) at 3;
    // an RLWE x RGSW step's forward transform at 1, its products at 0 and its
    // inverse (the last phase) at 2.  Measured at B = 4096 on one box: 114.0
    // ms without priorities, 111.3 with the automorphism step alone, 108.7
    // with the inverse at 2 as well, 107.3 with the forward at 1; the
    // automorphism step must rank above the inverse (at equal levels 111.8,
    // and 111.9 with the automorphism's own inverse at 2).  One bootstrap per
    // CU has nothing to arbitrate.
    constexpr bool PRIO = PZ_STEP_PRIO && R == 8;
    if (PRIO && is_auto) __builtin_amdgcn_s_setprio(3);
    // half 1 of an automorphism has no transform (its digits are not used)
    const bool work = !is_auto || h == 0;
    // the step's decomposer, field by field (a select of the whole struct is
    // built as a stack copy)
#define PZ_DEC_FIELD(f) (is_auto ? c.dec_auto.f : (h ? c.dec_b.f : c.dec_a.f))
    const u64 ca0 = PZ_DEC_FIELD(cadd0), ca1 = PZ_DEC_FIELD(cadd1), ca2 = PZ_DEC_FIELD(cadd2);
    const u32 dib = PZ_DEC_FIELD(ignore_bits);
#undef PZ_DEC_FIELD
    u64 D[1][R];
    // product rows: RLWE x RGSW -- own digits with the own-part row (half 0:
    // A(-sm), row 0; half 1: B(m), row 3), then the other half's digits with
    // the other row (A(m), row 2; B(-sm), row 1); automorphism -- one product
    // each, half 0 row A (0) with its digits, half 1 row B (1) with half 0's
    const size_t r1 = (size_t)(is_auto ? (h ? 1 : 0) : (h ? 3 : 0)) * N;
    const size_t r2 = (size_t)(h ? 1 : 2) * N;
    // unreduced products: at most two of them, so the accumulator stays below
    // 4q (8q with the [0,4q) products), the fast inverse's input bound
    constexpr bool Q4 = PZ_QUOT4 && PZ_LAZY_MAC && PZ_INV_FAST;
    // key rows loaded whole (32 VGPRs), the first before the barrier that
    // precedes its product (half 1 of an automorphism, with no transform to
    // run, before it waits out half 0's), the second before the read of the
    // other half's digits and the barrier after it: the load latencies
    // overlap the barrier waits instead of one exposed latency per half row
    KeySlice<R> k1, k2;
    constexpr int PF = R == 8 ? PZ_TW_PREFETCH : PZ_TW_PREFETCH4;
    PhaseTw<R> tw;  // the first phase's twiddles of the forward, then of the inverse
    constexpr int FMODE = PZ_QUOT4 ? (PZ_FWD6 ? FWD_FAST6 : FWD_FAST4) : PZ_FWD_MODE;
    if (work) {
        if (PRIO && !is_auto) __builtin_amdgcn_s_setprio(1);
        if constexpr (PF & 1) {
            fwd_tw_first<LOGN, R, FMODE>(tw, j, c.tb);
            __builtin_amdgcn_sched_barrier(0);  // issued ahead of the digits
        }
        // q + digit (below 2q): the unreduced forward transform takes it as is
#pragma unroll
        for (int r = 0; r < R; r++) D[0][r] = decomp_digit_d1<LOOSE>(q, dib, ca0, ca1, ca2, p[r]);
        if constexpr (PF & 1)
            ntt_fwd_pf<LOGN, 1, FMODE, PZ_SPLIT_SM, R, false>(D, j, mine, c.tb, tw);
        else
            ntt_fwd<LOGN, 1, FMODE, PZ_SPLIT_SM, R, false>(D, j, mine, c.tb);
#pragma unroll
        for (int r = 0; r < R; r++) mine[lds_ix(eval_x<LOGN, R>(j, r))] = D[0][r];
        // not hoisted into the transform: its twiddle loads would wait behind
        // the key row on the in-order load counter (issued after the last
        // phase's twiddles instead, measured no faster)
        __builtin_amdgcn_sched_barrier(0);
        load_slice(k1, K + r1, Ks + r1, j);
        if (PRIO && !is_auto) __builtin_amdgcn_s_setprio(0);
    } else {
        load_slice(k1, K + r1, Ks + r1, j);
        ntt_fwd_idle<LOGN, R>();
    }
    __syncthreads();
    if (!work) {
#pragma unroll
        for (int r = 0; r < R; r++) D[0][r] = theirs[lds_ix(eval_x<LOGN, R>(j, r))];
    }
    u64 acc[1][R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[0][r] = 0;
    fma_slice<PZ_SPLIT_SM, !PZ_LAZY_MAC, Q4>(acc[0], k1, D[0], q, q2);
    __builtin_amdgcn_sched_barrier(0);  // one key row in registers at a time
    if (!is_auto) {
        load_slice(k2, K + r2, Ks + r2, j);
#pragma unroll
        for (int r = 0; r < R; r++) D[0][r] = theirs[lds_ix(eval_x<LOGN, R>(j, r))];
    }
    // after the key row, which is needed first
    if constexpr ((PF & 6) == 2) inv_tw_first<LOGN, R, Q4>(tw, j, c.tb);
    __syncthreads();  // both halves have read: the regions may be rewritten
    if (!is_auto) fma_slice<PZ_SPLIT_SM, !PZ_LAZY_MAC, Q4>(acc[0], k2, D[0], q, q2);
    constexpr int IBND = Q4 ? 8 : (PZ_LAZY_MAC ? 4 : 2);  // accumulators below IBND q
    if (PRIO && !is_auto) __builtin_amdgcn_s_setprio(2);
    if constexpr ((PF & 6) == 4) inv_tw_first<LOGN, R, Q4>(tw, j, c.tb);
    if constexpr ((PF & 6) != 0)
        ntt_inv_pf<LOGN, 1, !LOOSE, false, PZ_SPLIT_SM, PZ_INV_FAST, IBND, R, false>(acc, j, mine,
                                                                                   c.tb, tw);
    else  // n^-1: key rows
        ntt_inv<LOGN, 1, !LOOSE, false, PZ_SPLIT_SM, PZ_INV_FAST, IBND, R, false>(acc, j, mine,
                                                                                c.tb);
    if (is_auto && h == 1) {
        // acc + (+-v) mod q, branch-free: the negated value q - v is selected,
        // then one conditional subtraction (acc < q, q - v <= q); the same word
        // as sub_mod / add_mod.  LOOSE: acc and v in [0, 2q), the negation
        // 2q - v, the result in [0, 2q)
#pragma unroll
        for (int r = 0; r < R; r++) {
            const u64 bp = stash[lds_ix(coef_x<LOGN, R>(j, r))];
            const u64 v = bp & ~SIGN;
            const u64 o = acc[0][r] + ((bp & SIGN) ? qn - v : v);
            p[r] = umin(o, o - qn);
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) p[r] = acc[0][r];
    }
    if (PRIO) __builtin_amdgcn_s_setprio(0);
    // the stash is rewritten only by a following automorphism, after its
    // leading barrier
}

}  // namespace pz
