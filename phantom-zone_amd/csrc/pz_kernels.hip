// pz_kernels.hip -- element-wise and row kernels of the bootstrapping path
// (gfx950). The RLWE operators and the blind rotation are in pz_rlwe.hip.
//
//   k_ntt_rows        batched Ntt::{forward,forward_lazy,backward,backward_lazy}
//   k_fma             ShoupMatrixFMA::shoup_matrix_fma
//   k_decompose       decompose_r (all digits)
//   k_prep            gate linear step + f64 moddown + LWE-KS digits
//   k_ks              LWE key switch (digits x KSK, mod lwe_q)
//   k_shoup           Shoup twins of key rows (ToShoup)
//   k_not / k_*       small element-wise helpers (NOT gate, keygen)
#include "pz_internal.hpp"

namespace pz {

// ---------------------------------------------------------------------------
// NTT rows: one work group (N/8 threads) per row.
// ---------------------------------------------------------------------------
template <int LOGN, int MODE>
__global__ __launch_bounds__(1 << (LOGN - 3)) void k_ntt_rows(u64 *rows, size_t nrows, Tables tb)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ u64 lds[];
    const u32 j = threadIdx.x;
    for (size_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        u64 *p = rows + row * N;
        u64 v[1][8];
        if (MODE <= 1) {
#pragma unroll
            for (int r = 0; r < 8; r++) v[0][r] = p[coef_x<LOGN>(j, r)];
            if (MODE == 0) ntt_fwd<LOGN, 1, FWD_LAZY>(v, j, lds, tb);
            else ntt_fwd<LOGN, 1, FWD_STRICT>(v, j, lds, tb);
#pragma unroll
            for (int r = 0; r < 8; r++) p[eval_x<LOGN>(j, r)] = v[0][r];
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) v[0][r] = p[eval_x<LOGN>(j, r)];
            if (MODE == 2) ntt_inv<LOGN, 1, false>(v, j, lds, tb);
            else ntt_inv<LOGN, 1, true>(v, j, lds, tb);
#pragma unroll
            for (int r = 0; r < 8; r++) p[coef_x<LOGN>(j, r)] = v[0][r];
        }
    }
}

// ---------------------------------------------------------------------------
// shoup_matrix_fma over d rows of length n (element-wise, grid-stride).
// ---------------------------------------------------------------------------
__global__ void k_fma(u64 *out, const u64 *a, const u64 *as, const u64 *b, u32 d, size_t n,
                      u64 q)
{
    const u64 q2 = q << 1;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 o = out[i];
        for (u32 r = 0; r < d; r++) {
            size_t k = (size_t)r * n + i;
            o = shoup_fma(o, a[k], as[k], b[k], q, q2);
        }
        out[i] = o;
    }
}

// decompose_r: digits[r][i] = digit r of in[i]
__global__ void k_decompose(const u64 *in, u64 *digits, size_t count, Decomp dc)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 value = decomp_prepare(dc, in[i]);
        for (u32 r = 0; r < dc.d; r++) digits[(size_t)r * count + i] = decomp_next(dc, value);
    }
}

// ---------------------------------------------------------------------------
// k_prep: per ciphertext coefficient -- gate linear step (src/bool/
// evaluator.rs:2118-2133), f64 moddown Q -> q_ks (src/pbs.rs:128-132) and
// the LWE-KS decomposition (src/lwe.rs:31-35) into signed int8 digits.
// mode: PREP_GATE (c0, c1, ops), PREP_PBS (in), PREP_KSONLY (in is mod q_ks).
// One work group per tile of KS_CT ciphertexts x PREP_COLS mask columns: the
// inputs are read along rows (8 consecutive words per ciphertext), the digits
// are assembled in LDS in the key switch's layout [tile][column][digit][KS_CT]
// and leave as one contiguous run of PREP_COLS * d * KS_CT bytes.
// ---------------------------------------------------------------------------
constexpr int PREP_COLS = 64;
constexpr int PREP_MAXD = 17;  // lwe_logb * lwe_d <= log2(q_ks) <= 17

PZ_DEV u64 prep_value(int mode, const u64 *c0, const u64 *c1, size_t r0, size_t r1, u32 op,
                      u32 col, u32 N1, const DevCtx &c)
{
    const u64 q = c.tb.q;
    u64 v = c0[r0 * N1 + col];
    if (mode == PREP_GATE) {
        const u64 v1 = c1[r1 * N1 + col];
        if (op <= 3) {
            v = add_mod(v, v1, q);
            if (col == 0) v = add_mod(v, c.qby4, q);
        } else {
            v = sub_mod(v, v1, q);
            v = add_mod(v, v, q);
        }
    }
    if (mode != PREP_KSONLY) {
        const double x = ((double)v * c.lwe_qf) / c.rlwe_qf;
        v = (u64)round(x);
    }
    return v;
}

__global__ __launch_bounds__(256) void k_prep(int mode, const u64 *c0, const u64 *c1,
                                              const u32 *idx0, const u32 *idx1, const u32 *ops,
                                              u32 N1, size_t B, int8_t *digits, u32 *ks_b,
                                              u32 *ks_out, DevCtx c, bool rowmajor)
{
    constexpr int CW = KS_CT / 4;  // u32 words of one digit row of a tile
    __shared__ u32 sd[PREP_COLS * PREP_MAXD * CW];
    const u32 tid = threadIdx.x;
    const size_t b0 = (size_t)blockIdx.x * KS_CT;
    const u32 cb = blockIdx.y * PREP_COLS;  // first mask column (0-based) of this group
    const u32 d = c.lwe_d, n1 = c.n_lwe + 1;
    const u64 lq = c.lwe_mask + 1;
    // this group's share of the key-switch output, cleared (k_ks accumulates)
    for (u32 e = tid; e < KS_CT * PREP_COLS; e += blockDim.x) {
        const size_t bb = b0 + e / PREP_COLS;
        const u32 col = cb + e % PREP_COLS;
        if (bb < B && col < n1) ks_out[bb * n1 + col] = 0;
    }
    // body coefficient (column 0) of each ciphertext
    if (blockIdx.y == 0 && tid < KS_CT && b0 + tid < B) {
        const size_t bb = b0 + tid;
        ks_b[bb] = (u32)prep_value(mode, c0, c1, idx0 ? idx0[bb] : bb, idx1 ? idx1[bb] : bb,
                                   mode == PREP_GATE ? ops[bb] : 0, 0, N1, c);
    }
    // digits: item (m, cl) takes ciphertexts 4m..4m+3 of the tile in column
    // 1 + cb + cl and packs their digit bytes; ciphertexts past B give 0
    static_assert(PREP_COLS * CW % 256 == 0, "whole rounds of the 256 threads");
#pragma unroll
    for (u32 e0 = 0; e0 < PREP_COLS * CW; e0 += 256) {
        const u32 m = (e0 + tid) % CW, cl = (e0 + tid) / CW;
        const u32 col = 1 + cb + cl;
        u32 w[PREP_MAXD];
#pragma unroll
        for (int r = 0; r < PREP_MAXD; r++) w[r] = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const size_t bb = b0 + 4 * m + i;
            if (bb >= B) continue;
            const u64 v = prep_value(mode, c0, c1, idx0 ? idx0[bb] : bb, idx1 ? idx1[bb] : bb,
                                     mode == PREP_GATE ? ops[bb] : 0, col, N1, c);
            u64 value = decomp_prepare(c.dec_lwe, v);
#pragma unroll
            for (int r = 0; r < PREP_MAXD; r++) {
                if ((u32)r >= d) break;
                const u64 dg = decomp_next(c.dec_lwe, value);
                const int sv = dg >= (lq >> 1) ? (int)((long long)dg - (long long)lq) : (int)dg;
                w[r] |= (u32)(uint8_t)(int8_t)sv << (8 * i);
            }
        }
#pragma unroll
        for (int r = 0; r < PREP_MAXD; r++) {
            if ((u32)r >= d) break;
            sd[(cl * d + r) * CW + m] = w[r];
        }
    }
    __syncthreads();
    const size_t rows = (size_t)(N1 - 1) * d;
    if (rowmajor) {
        // [ciphertext][row]: a ciphertext's digits of this group's columns are one
        // run of PREP_COLS * d bytes from row cb * d; each item gathers 16
        // consecutive rows of one ciphertext from the LDS words (byte ct % 4 of
        // word ct / 4 of each row).  The buffer holds whole tiles of KS_CT, and
        // ciphertexts past B carry zero digits.
        const u32 chunks = PREP_COLS * d / 16;
        for (u32 e = tid; e < KS_CT * chunks; e += blockDim.x) {
            const u32 ct = e / chunks, pos0 = (e % chunks) * 16;
            const u32 sh = 8 * (ct % 4);
            u32 wv[4];
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++) {
                u32 word = 0;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    word |= ((sd[(pos0 + 4 * q4 + i) * CW + ct / 4] >> sh) & 0xffu) << (8 * i);
                wv[q4] = word;
            }
            *reinterpret_cast<uint4 *>(digits + (b0 + ct) * rows + (size_t)cb * d + pos0) =
                make_uint4(wv[0], wv[1], wv[2], wv[3]);
        }
        return;
    }
    // the group's columns are one contiguous run of the tile
    uint4 *dst = reinterpret_cast<uint4 *>(digits + (b0 / KS_CT) * rows * KS_CT +
                                           (size_t)cb * d * KS_CT);
    const uint4 *src = reinterpret_cast<const uint4 *>(sd);
    for (u32 e = tid; e < PREP_COLS * d * CW / 4; e += blockDim.x) dst[e] = src[e];
}

// ---------------------------------------------------------------------------
// k_ks: LWE key switch as a digits x KSK product, mod q_ks (src/lwe.rs:14-44).
//   out[b][col] += sum_{rows in this split} digit[b][row] * ksk[row][col]
// grid (ciphertext tiles, row splits); each thread owns one output unit for KS_CT
// ciphertexts: for q_ks <= 2^16 a unit is two adjacent columns held as 16-bit
// lanes of one register (v_pk_mad_u16: both lanes wrap mod 2^16, which is
// the reference's mask arithmetic); otherwise one column in 32 bits.  Digits
// are staged through LDS, pre-broadcast to both lanes.  The row splits meet
// in u32 atomics; readers mask the sums with q_ks - 1 (q_ks divides 2^32).
// out must be zero on entry (k_prep clears it) and gets b_in at column 0.
// ---------------------------------------------------------------------------
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Ciphertexts accumulated per work group: the whole KS_CT tile with 16-bit
// lanes, at most 32 with 32-bit accumulators (register budget); ks_subs work
// groups share a tile.  (Tiles of 64 ran slower than 32: the accumulators
// cost occupancy, and the key re-reads they save are not the bound.)
template <bool PACK16>
constexpr int ks_ct() { return PACK16 || KS_CT <= 32 ? KS_CT : KS_CT / 2; }
template <bool PACK16>
constexpr int ks_subs() { return KS_CT / ks_ct<PACK16>(); }

template <bool PACK16>
__global__ __launch_bounds__(1024) void k_ks(const int8_t *digits, const u32 *ks_b,
                                             const void *ksk, u32 rows, u32 n1, u32 ldk,
                                             size_t seg_begin, size_t seg_end, u32 *out,
                                             const u32 *work, const DevKey *keys)
{
    constexpr int CT = ks_ct<PACK16>(), SUBS = ks_subs<PACK16>();
    __shared__ __attribute__((aligned(16))) u32 dig[KS_ROWS][CT];
    const u32 unit = threadIdx.x;
    const u32 units = PACK16 ? (n1 + 1) / 2 : n1;
    // this group's CT ciphertexts (part `sub` of a tile of KS_CT); only those
    // in [seg_begin, seg_end) (the ciphertexts of this key) are accumulated.
    // With a work list, entry blockIdx.x / SUBS names the tile, the segment
    // and the key (batches that mix server keys: one launch for every key).
    const u32 sub = blockIdx.x % SUBS, gx = blockIdx.x / SUBS;
    size_t tile_id = seg_begin / KS_CT + gx;
    if (work) {
        const u32 *e = work + 4 * (size_t)gx;
        tile_id = e[0];
        seg_begin = e[1];
        seg_end = e[2];
        ksk = keys[e[3]].ksk;
    }
    const size_t b0 = tile_id * KS_CT + sub * CT;
    const int8_t *tile = digits + tile_id * rows * KS_CT + sub * CT;
    const u32 per = (rows + gridDim.y - 1) / gridDim.y;
    const u32 rbeg = blockIdx.y * per;
    const u32 rend = rbeg + per < rows ? rbeg + per : rows;
    const uint32_t *k32 = reinterpret_cast<const uint32_t *>(ksk);
    const u32 ld32 = PACK16 ? ldk / 2 : ldk;
    u32 acc[CT];
#pragma unroll
    for (int t = 0; t < CT; t++) acc[t] = 0;
    for (u32 r0 = rbeg; r0 < rend; r0 += KS_ROWS) {
        const u32 rmax = (rend - r0) < (u32)KS_ROWS ? (rend - r0) : (u32)KS_ROWS;
        __syncthreads();
        // stage: 4 digits (one u32 of int8) per step, broadcast to 16-bit lanes
        for (u32 e = threadIdx.x; e < rmax * CT / 4; e += blockDim.x) {
            const u32 rr = e / (CT / 4), c4 = e % (CT / 4);
            u32 w4 = *reinterpret_cast<const uint32_t *>(tile + (size_t)(r0 + rr) * KS_CT + 4 * c4);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int dv = (int)(int8_t)(w4 >> (8 * k));
                (&dig[0][0])[4 * e + k] = PACK16 ? ((u32)(uint16_t)(int16_t)dv) * 0x10001u : (u32)dv;
            }
        }
        __syncthreads();
        if (unit < units) {
            const uint32_t *kp = k32 + (size_t)r0 * ld32 + unit;
            u32 rr = 0;
            for (; rr + 4 <= rmax; rr += 4) {
                u32 kv[4];
#pragma unroll
                for (int u = 0; u < 4; u++) kv[u] = kp[(size_t)(rr + u) * ld32];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const uint4 *dr = reinterpret_cast<const uint4 *>(&dig[rr + u][0]);
#pragma unroll
                    for (int t4 = 0; t4 < CT / 4; t4++) {
                        uint4 dd = dr[t4];
                        u32 dv[4] = {dd.x, dd.y, dd.z, dd.w};
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            int t = 4 * t4 + k;
                            if (PACK16) {
                                u16x2 a = __builtin_bit_cast(u16x2, acc[t]);
                                u16x2 dk = __builtin_bit_cast(u16x2, dv[k]);
                                u16x2 kk = __builtin_bit_cast(u16x2, kv[u]);
                                acc[t] = __builtin_bit_cast(u32, (u16x2)(dk * kk + a));
                            } else {
                                acc[t] += (u32)((int)dv[k] * (int)kv[u]);
                            }
                        }
                    }
                }
            }
            for (; rr < rmax; rr++) {
                u32 kv = kp[(size_t)rr * ld32];
#pragma unroll
                for (int t = 0; t < CT; t++) {
                    if (PACK16) {
                        u16x2 a = __builtin_bit_cast(u16x2, acc[t]);
                        u16x2 dk = __builtin_bit_cast(u16x2, dig[rr][t]);
                        u16x2 kk = __builtin_bit_cast(u16x2, kv);
                        acc[t] = __builtin_bit_cast(u32, (u16x2)(dk * kk + a));
                    } else {
                        acc[t] += (u32)((int)dig[rr][t] * (int)kv);
                    }
                }
            }
        }
    }
    if (unit >= units) return;
#pragma unroll
    for (int t = 0; t < CT; t++) {
        size_t bb = b0 + t;
        if (bb < seg_begin || bb >= seg_end) continue;
        u32 *o = out + bb * n1;
        if (PACK16) {
            u32 c0 = 2 * unit;
            u32 lo = acc[t] & 0xffffu, hi = acc[t] >> 16;
            if (c0 == 0 && blockIdx.y == 0) lo += ks_b[bb];
            atomicAdd(o + c0, lo);
            if (c0 + 1 < n1) atomicAdd(o + c0 + 1, hi);
        } else {
            u32 v = acc[t];
            if (unit == 0 && blockIdx.y == 0) v += ks_b[bb];
            atomicAdd(o + unit, v);
        }
    }
}

// ---------------------------------------------------------------------------
// k_ks_mfma: the LWE key switch (src/lwe.rs:14-44) as an int8 matrix product on
// the matrix cores.  out[b][col] += sum_row digit[b][row] * ksk[row][col] is an
// exact integer contraction mod q_ks: the digits are signed and small
// (|digit| <= 2^(logb-1) <= 8) and each KSK word is split into KS_PLANES 7-bit
// planes, so every product of v_mfma_i32_16x16x64_i8 is exact and every int32
// sum stays below 2^31 (rows * 8 * 127 < 2^31 for rows <= 2^21); the planes
// recombine as P0 + P1 2^7 + P2 2^14 in u32 arithmetic, whose wrap is the
// reference's mask arithmetic mod q_ks (q_ks divides 2^32).
// Operands: digits row-major [ciphertext][row] (k_prep, rowmajor), planes
// [plane][ldk][row] (launch_pack_ksk_planes); both keep 16 consecutive rows of
// one ciphertext / column together, the operand a lane of the 16x16x64 form
// holds (lane l: ciphertext or column l % 16, rows 16 (l / 16) .. + 16; output
// slot i of lane l: ciphertext 4 (l / 16) + i, column l % 16).
// Work group: KSM_WM x 2 waves, each 32 ciphertexts x 48 columns; a work group
// covers KSM_WM * 32 ciphertexts x 96 columns over one row range (blockIdx.z;
// about PZ_KSM_GROUPS work groups per launch).  Measured (one box): 4096 NI_2P
// ciphertexts 0.32 ms (8 waves x 2: 0.33; 256 or 1024 groups: 0.38 / 0.34), 512
// NI_8P ciphertexts 0.08 ms, against 1.63 / 0.66 ms for k_ks;
// each 64-row block of both operands is staged through LDS (rows padded to 80
// bytes: the 16 lanes of a b128 read start in distinct bank groups) and the
// next block's loads are issued before the current block's products.  Row
// ranges meet in u32 atomics on the zeroed output, like k_ks.
// ---------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
#ifndef PZ_KSM_GROUPS
#define PZ_KSM_GROUPS 512
#endif
#ifndef PZ_KSM_WM
#define PZ_KSM_WM 4
#endif
constexpr int KSM_WM = PZ_KSM_WM;         // waves along the ciphertexts
constexpr int KSM_TM = 32 * KSM_WM;       // ciphertexts per work group
constexpr int KSM_TN = 96;                // columns per work group
constexpr int KSM_KB = 64;                // rows per block
constexpr int KSM_LD = 80;                // LDS pitch of one ciphertext / column
constexpr int KSM_THREADS = 128 * KSM_WM;
constexpr int KSM_BCH = KS_PLANES * KSM_TN * (KSM_KB / 16);  // 16-byte chunks of B per block
constexpr int KSM_BPT = (KSM_BCH + KSM_THREADS - 1) / KSM_THREADS;

__global__ __launch_bounds__(KSM_THREADS) void k_ks_mfma(const int8_t *digits, const u32 *ks_b,
                                                       const int8_t *planes, u32 rows, u32 n1,
                                                       u32 ldk, size_t seg_begin, size_t B,
                                                       size_t Bpad, u32 kchunk, u32 *out)
{
    __shared__ __attribute__((aligned(16))) int8_t sa[KSM_TM * KSM_LD];
    __shared__ __attribute__((aligned(16))) int8_t sb[KS_PLANES * KSM_TN * KSM_LD];
    const u32 tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const u32 wm = w % KSM_WM, wn = w / KSM_WM;
    const size_t ct0 = (size_t)blockIdx.x * KSM_TM;
    const u32 col0 = blockIdx.y * KSM_TN;
    const u32 kbeg = blockIdx.z * kchunk;
    const u32 kend = kbeg + kchunk < rows ? kbeg + kchunk : rows;
    // staging: A chunk (ciphertext tid / 4, 16 rows at 16 (tid % 4)); B chunks e
    const size_t a_ct = ct0 + tid / 4;
    const bool a_ok = a_ct < Bpad;
    const int8_t *a_src = digits + (a_ct < Bpad ? a_ct : 0) * rows + 16 * (tid % 4);
    v4i ra, rb[KSM_BPT];
    const auto load = [&](u32 k0) {
        ra = a_ok ? *reinterpret_cast<const v4i *>(a_src + k0) : v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < KSM_BPT; i++) {
            const u32 e = tid + i * KSM_THREADS;
            const u32 pl = e / (KSM_TN * 4), rem = e % (KSM_TN * 4), col = col0 + rem / 4;
            rb[i] = (e < (u32)KSM_BCH && col < ldk)
                        ? *reinterpret_cast<const v4i *>(planes + ((size_t)pl * ldk + col) * rows +
                                                         k0 + 16 * (rem % 4))
                        : v4i{0, 0, 0, 0};
        }
    };
    v4i acc[KS_PLANES][2][3];
#pragma unroll
    for (int p = 0; p < KS_PLANES; p++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) acc[p][i][j] = v4i{0, 0, 0, 0};
    if (kbeg < kend) load(kbeg);
    for (u32 k0 = kbeg; k0 < kend; k0 += KSM_KB) {
        __syncthreads();  // the previous block's fragment reads are done
        *reinterpret_cast<v4i *>(sa + (tid / 4) * KSM_LD + 16 * (tid % 4)) = ra;
#pragma unroll
        for (int i = 0; i < KSM_BPT; i++) {
            const u32 e = tid + i * KSM_THREADS;
            if (e < (u32)KSM_BCH) {
                const u32 pl = e / (KSM_TN * 4), rem = e % (KSM_TN * 4);
                *reinterpret_cast<v4i *>(sb + (pl * KSM_TN + rem / 4) * KSM_LD + 16 * (rem % 4)) =
                    rb[i];
            }
        }
        __syncthreads();
        if (k0 + KSM_KB < kend) load(k0 + KSM_KB);  // in flight during the products
        v4i af[2];
#pragma unroll
        for (int i = 0; i < 2; i++)
            af[i] = *reinterpret_cast<const v4i *>(sa + (32 * wm + 16 * i + lane % 16) * KSM_LD +
                                                   16 * (lane / 16));
#pragma unroll
        for (int p = 0; p < KS_PLANES; p++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const v4i bf = *reinterpret_cast<const v4i *>(
                    sb + (p * KSM_TN + 48 * wn + 16 * j + lane % 16) * KSM_LD + 16 * (lane / 16));
#pragma unroll
                for (int i = 0; i < 2; i++)
                    acc[p][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf, acc[p][i][j],
                                                                         0, 0, 0);
            }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int sl = 0; sl < 4; sl++) {
                const size_t ct = ct0 + 32 * wm + 16 * i + 4 * (lane / 16) + sl;
                const u32 col = col0 + 48 * wn + 16 * j + lane % 16;
                if (ct < seg_begin || ct >= B || col >= n1) continue;
                u32 v = 0;
#pragma unroll
                for (int p = 0; p < KS_PLANES; p++) v += (u32)acc[p][i][j][sl] << (7 * p);
                if (col == 0 && blockIdx.z == 0) v += ks_b[ct];
                atomicAdd(out + ct * n1 + col, v);
            }
}

// KSK rows [rows][n1] u64 (below q_ks <= 2^21) -> 7-bit planes [plane][ldk][rows]
__global__ void k_pack_ksk_planes(const u64 *in, int8_t *out, size_t rows, u32 n1, u32 ldk)
{
    const size_t n = (size_t)ldk * rows;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const u32 col = (u32)(i / rows);
        const size_t r = i % rows;
        const u64 v = col < n1 ? in[r * n1 + col] : 0;
#pragma unroll
        for (int p = 0; p < KS_PLANES; p++) out[(size_t)p * n + i] = (int8_t)((v >> (7 * p)) & 127u);
    }
}

// Shoup twins: out[i] = floor(in[i] * 2^64 / q) (ToShoup, src/utils.rs:49-73)
// by restoring long division of (v << 64) by q, v < q.
// x[i] = x[i] * c mod q in [0,q) (c < q with Shoup twin c_s; x[i] < q)
__global__ void k_scale(u64 *x, size_t n, u64 c, u64 c_s, u64 q)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 v = shoup_mul(x[i], c, c_s, q);
        x[i] = umin(v, v - q);
    }
}

// Shoup twins floor(w*2^64/q) (ToShoup, src/utils.rs:49-73) for w < q.  With
// 2^64 = Q0*q + R0 the twin is w*Q0 + floor(w*R0/q); the last quotient is the
// Shoup estimate hi64(w*R0s), R0s = floor(R0*2^64/q), whose remainder
// w*R0 - q*hi64(w*R0s) lies in [0, 2q): one correction step makes it exact.
__global__ void k_shoup(const u64 *in, u64 *out, size_t n, u64 q, u64 Q0, u64 R0, u64 R0s)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const u64 w = in[i];
        const u64 qt = __umul64hi(w, R0s);
        const u64 r = w * R0 - q * qt;
        out[i] = w * Q0 + qt + (r >= q ? 1 : 0);
    }
}

// NOT gate: every element negated (src/bool/evaluator.rs:2236-2239)
__global__ void k_not(u64 *c0, size_t n, u64 q)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        c0[i] = neg_mod(c0[i], q);
}

// ---------------------------------------------------------------------------
// k_valu_probe: issue-rate microbenchmark of the VALU instruction forms the
// PBS kernels are built from.  Each thread runs PROBE_CH independent chains
// for `iters` steps (an empty asm keeps every step live) and stores the xor
// of its chains.  kind: 0 v_add_u32, 1 v_mul_lo_u32, 2 v_mul_hi_u32,
// 3 v_mad_u64_u32, 4 one 54-bit Shoup multiply (shoup_mul), 5 v_lshl_add_u64
// (the butterflies' 64-bit 2x + c).
// ---------------------------------------------------------------------------
constexpr int PROBE_CH = 8;

template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(u64 *out, u32 iters, u32 c, u64 w, u64 ws,
                                                    u64 q)
{
    const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    u64 acc = 0;
    if (KIND >= 3) {
        u64 x[PROBE_CH];
#pragma unroll
        for (int k = 0; k < PROBE_CH; k++) x[k] = gid * 0x9e3779b97f4a7c15ull + (u64)k;
        for (u32 i = 0; i < iters; i++) {
#pragma unroll
            for (int k = 0; k < PROBE_CH; k++) {
                if (KIND == 3) x[k] = (u64)(u32)x[k] * c + x[k];
                else if (KIND == 5) x[k] = twice_plus(x[k], w);
                else x[k] = shoup_mul(x[k], w, ws, q);
                asm volatile("" : "+v"(x[k]));
            }
        }
#pragma unroll
        for (int k = 0; k < PROBE_CH; k++) acc ^= x[k];
    } else {
        u32 y[PROBE_CH];
#pragma unroll
        for (int k = 0; k < PROBE_CH; k++) y[k] = (u32)gid * 2654435761u + (u32)k;
        for (u32 i = 0; i < iters; i++) {
#pragma unroll
            for (int k = 0; k < PROBE_CH; k++) {
                if (KIND == 0) y[k] = y[k] + c;
                else if (KIND == 1) y[k] = y[k] * c;
                else y[k] = __umulhi(y[k], c);
                asm volatile("" : "+v"(y[k]));
            }
        }
#pragma unroll
        for (int k = 0; k < PROBE_CH; k++) acc ^= y[k];
    }
    out[gid] = acc;
}

void launch_valu_probe(int kind, int grid, u64 *out, u32 iters, u32 c, u64 w, u64 ws, u64 q,
                       hipStream_t s)
{
    switch (kind) {
    case 0: hipLaunchKernelGGL(k_valu_probe<0>, grid, 256, 0, s, out, iters, c, w, ws, q); break;
    case 1: hipLaunchKernelGGL(k_valu_probe<1>, grid, 256, 0, s, out, iters, c, w, ws, q); break;
    case 2: hipLaunchKernelGGL(k_valu_probe<2>, grid, 256, 0, s, out, iters, c, w, ws, q); break;
    case 3: hipLaunchKernelGGL(k_valu_probe<3>, grid, 256, 0, s, out, iters, c, w, ws, q); break;
    case 5: hipLaunchKernelGGL(k_valu_probe<5>, grid, 256, 0, s, out, iters, c, w, ws, q); break;
    default: hipLaunchKernelGGL(k_valu_probe<4>, grid, 256, 0, s, out, iters, c, w, ws, q); break;
    }
}

// u32 LWE words -> u64
__global__ void k_widen(const u32 *in, u64 *out, size_t n, u32 mask)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i] & mask;
}

// KSK rows (u64, reference layout [rows][n1]) -> device words [rows][ldk]
// (ldk >= n1, padding zero)
template <typename T>
__global__ void k_pack_ksk(const u64 *in, T *out, size_t rows, u32 n1, u32 ldk)
{
    size_t n = rows * ldk;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t r = i / ldk;
        u32 col = (u32)(i % ldk);
        out[i] = col < n1 ? (T)in[r * n1 + col] : (T)0;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int grid_for(size_t n, int tpb, int cap = 65535 * 8)
{
    size_t g = (n + tpb - 1) / tpb;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

template <int LOGN>
static void launch_ntt_t(int mode, u64 *rows, size_t nrows, const Tables &tb, hipStream_t s)
{
    constexpr int TPB = 1 << (LOGN - 3);
    size_t lds = 8 * (size_t)Geom<LOGN>::LDS_WORDS;
    int grid = nrows < 65535 ? (int)nrows : 65535;
    if (grid < 1) return;
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_ntt_rows<LOGN, 0>), grid, TPB, lds, s, rows, nrows, tb); break;
    case 1: hipLaunchKernelGGL((k_ntt_rows<LOGN, 1>), grid, TPB, lds, s, rows, nrows, tb); break;
    case 2: hipLaunchKernelGGL((k_ntt_rows<LOGN, 2>), grid, TPB, lds, s, rows, nrows, tb); break;
    default: hipLaunchKernelGGL((k_ntt_rows<LOGN, 3>), grid, TPB, lds, s, rows, nrows, tb); break;
    }
}

void launch_ntt(int logn, int mode, u64 *rows, size_t nrows, const Tables &tb, hipStream_t s)
{
    switch (logn) {
    case 9: launch_ntt_t<9>(mode, rows, nrows, tb, s); break;
    case 11: launch_ntt_t<11>(mode, rows, nrows, tb, s); break;
    }
}

void launch_fma(u64 *out, const u64 *a, const u64 *as, const u64 *b, u32 d, size_t n, u64 q,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_fma, grid_for(n, 256), 256, 0, s, out, a, as, b, d, n, q);
}

void launch_decompose(const u64 *in, u64 *digits, size_t count, const Decomp &dc, hipStream_t s)
{
    hipLaunchKernelGGL(k_decompose, grid_for(count, 256), 256, 0, s, in, digits, count, dc);
}

void launch_prep(int mode, const u64 *c0, const u64 *c1, const u32 *idx0, const u32 *idx1,
                 const u32 *ops, u32 N1, size_t B, int8_t *digits, u32 *ks_b, u32 *ks_out,
                 const DevCtx &c, hipStream_t s, bool rowmajor)
{
    if (B == 0) return;
    // (N1 - 1) is a multiple of PREP_COLS for every supported ring size
    dim3 grid((unsigned)((B + KS_CT - 1) / KS_CT), (N1 - 1) / PREP_COLS);
    hipLaunchKernelGGL(k_prep, grid, 256, 0, s, mode, c0, c1, idx0, idx1, ops, N1, B, digits,
                       ks_b, ks_out, c, rowmajor);
}

void launch_ks(const int8_t *digits, const u32 *ks_b, const void *ksk, u32 ksk_bytes, u32 rows,
               u32 n1, u32 ldk, size_t seg_begin, size_t seg_end, u32 *out, hipStream_t s,
               const int8_t *planes)
{
    if (seg_end <= seg_begin) return;
    if (PZ_KS_MFMA && planes) {
        // the grid covers ciphertexts [0, seg_end); outputs below seg_begin are
        // left alone.  rows is a multiple of KSM_KB (N * d with N >= 512).
        const size_t B = seg_end, Bpad = (B + KS_CT - 1) / KS_CT * KS_CT;
        const u32 gx = (u32)((Bpad + KSM_TM - 1) / KSM_TM), gy = (ldk + KSM_TN - 1) / KSM_TN;
        // row ranges so that the launch has about PZ_KSM_GROUPS work groups
        u32 split = (PZ_KSM_GROUPS + gx * gy - 1) / (gx * gy);
        const u32 blocks = rows / KSM_KB;
        if (split > blocks) split = blocks;
        if (split < 1) split = 1;
        const u32 kchunk = (blocks + split - 1) / split * KSM_KB;
        split = (rows + kchunk - 1) / kchunk;
        hipLaunchKernelGGL(k_ks_mfma, dim3(gx, gy, split), KSM_THREADS, 0, s, digits, ks_b, planes,
                           rows, n1, ldk, seg_begin, B, Bpad, kchunk, out);
        return;
    }
    const bool pack = ksk_bytes == 2;
    u32 units = pack ? (n1 + 1) / 2 : n1;
    u32 tpb = (units + 63) / 64 * 64;
    size_t tiles = (seg_end - 1) / KS_CT - seg_begin / KS_CT + 1;
    // split the rows so that even one tile fills the chip: about KS_GROUPS
    // work groups, at least KS_SPLIT row ranges, each at least KS_ROWS rows
    size_t split = (KS_GROUPS + tiles - 1) / tiles;
    size_t max_split = rows / KS_ROWS > 0 ? rows / KS_ROWS : 1;
    if (split > max_split) split = max_split;
    if (split < KS_SPLIT) split = KS_SPLIT;
    dim3 grid((unsigned)(tiles * (pack ? ks_subs<true>() : ks_subs<false>())), (unsigned)split);
    if (pack)
        hipLaunchKernelGGL(k_ks<true>, grid, tpb, 0, s, digits, ks_b, ksk, rows, n1, ldk,
                           seg_begin, seg_end, out, nullptr, nullptr);
    else
        hipLaunchKernelGGL(k_ks<false>, grid, tpb, 0, s, digits, ks_b, ksk, rows, n1, ldk,
                           seg_begin, seg_end, out, nullptr, nullptr);
}

// One key-switch launch over a work list of n_work (tile, segment begin,
// segment end, key index) entries; keys[i].ksk are KSKs of one layout.
void launch_ks_multi(const int8_t *digits, const u32 *ks_b, const DevKey *keys, const u32 *work,
                     size_t n_work, u32 ksk_bytes, u32 rows, u32 n1, u32 ldk, u32 *out,
                     hipStream_t s)
{
    if (n_work == 0) return;
    const bool pack = ksk_bytes == 2;
    u32 units = pack ? (n1 + 1) / 2 : n1;
    u32 tpb = (units + 63) / 64 * 64;
    size_t split = (KS_GROUPS + n_work - 1) / n_work;
    size_t max_split = rows / KS_ROWS > 0 ? rows / KS_ROWS : 1;
    if (split > max_split) split = max_split;
    if (split < KS_SPLIT) split = KS_SPLIT;
    dim3 grid((unsigned)(n_work * (pack ? ks_subs<true>() : ks_subs<false>())), (unsigned)split);
    if (pack)
        hipLaunchKernelGGL(k_ks<true>, grid, tpb, 0, s, digits, ks_b, nullptr, rows, n1, ldk,
                           (size_t)0, (size_t)0, out, work, keys);
    else
        hipLaunchKernelGGL(k_ks<false>, grid, tpb, 0, s, digits, ks_b, nullptr, rows, n1, ldk,
                           (size_t)0, (size_t)0, out, work, keys);
}

void launch_scale(u64 *x, size_t n, u64 c, u64 c_s, u64 q, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale, grid_for(n, 256), 256, 0, s, x, n, c, c_s, q);
}

void launch_shoup(const u64 *in, u64 *out, size_t n, u64 q, hipStream_t s)
{
    typedef unsigned __int128 u128;
    const u128 two64 = (u128)1 << 64;
    const u64 Q0 = (u64)(two64 / q), R0 = (u64)(two64 % q);
    const u64 R0s = (u64)(((u128)R0 << 64) / q);
    hipLaunchKernelGGL(k_shoup, grid_for(n, 256), 256, 0, s, in, out, n, q, Q0, R0, R0s);
}

void launch_not(u64 *c0, size_t n, u64 q, hipStream_t s)
{
    hipLaunchKernelGGL(k_not, grid_for(n, 256), 256, 0, s, c0, n, q);
}

void launch_widen(const u32 *in, u64 *out, size_t n, u32 mask, hipStream_t s)
{
    hipLaunchKernelGGL(k_widen, grid_for(n, 256), 256, 0, s, in, out, n, mask);
}

void launch_pack_ksk_planes(const u64 *in, int8_t *out, size_t rows, u32 n1, u32 ldk,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_ksk_planes, grid_for((size_t)ldk * rows, 256), 256, 0, s, in, out,
                       rows, n1, ldk);
}

void launch_pack_ksk(const u64 *in, void *out, size_t rows, u32 n1, u32 ldk, u32 bytes,
                     hipStream_t s)
{
    size_t n = rows * ldk;
    if (bytes == 2)
        hipLaunchKernelGGL(k_pack_ksk<uint16_t>, grid_for(n, 256), 256, 0, s, in,
                           (uint16_t *)out, rows, n1, ldk);
    else
        hipLaunchKernelGGL(k_pack_ksk<uint32_t>, grid_for(n, 256), 256, 0, s, in,
                           (uint32_t *)out, rows, n1, ldk);
}

// ---------------------------------------------------------------------------
// Seed expansion: rand 0.8.5 UniformInt<u64> draws (widening multiply,
// rejection zone; src/random.rs:93-110) from ChaCha8 streams (rand_chacha's
// layout: 64-bit block counter in words 12-13, stream 0 in 14-15).  One
// thread per 64-byte block, i.e. eight draws; a draw outside the zone marks
// its stream for sequential regeneration on the host.
// ---------------------------------------------------------------------------
PZ_DEV u32 rotl32(u32 v, int c) { return (v << c) | (v >> (32 - c)); }
#define PZ_QR(a, b, c, d)                 \
    a += b; d ^= a; d = rotl32(d, 16);    \
    c += d; b ^= c; b = rotl32(b, 12);    \
    a += b; d ^= a; d = rotl32(d, 8);     \
    c += d; b ^= c; b = rotl32(b, 7);

__global__ void k_expand_uniform(const u32 *keys, u32 S, u64 draws, u64 keep_from, u64 q,
                                 u64 zone, u64 *out, u32 *reject)
{
    const u64 blocks = (draws + 7) / 8;
    const u64 kept = draws - keep_from;
    for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < (u64)S * blocks;
         t += (u64)gridDim.x * blockDim.x) {
        const u32 s = (u32)(t / blocks);
        const u64 b = t % blocks;
        u32 st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; i++) st[4 + i] = keys[8 * (size_t)s + i];
        st[12] = (u32)b;
        st[13] = (u32)(b >> 32);
        st[14] = 0;
        st[15] = 0;
        u32 x[16];
        for (int i = 0; i < 16; i++) x[i] = st[i];
        for (int r = 0; r < 8; r += 2) {
            PZ_QR(x[0], x[4], x[8], x[12]);
            PZ_QR(x[1], x[5], x[9], x[13]);
            PZ_QR(x[2], x[6], x[10], x[14]);
            PZ_QR(x[3], x[7], x[11], x[15]);
            PZ_QR(x[0], x[5], x[10], x[15]);
            PZ_QR(x[1], x[6], x[11], x[12]);
            PZ_QR(x[2], x[7], x[8], x[13]);
            PZ_QR(x[3], x[4], x[9], x[14]);
        }
        bool bad = false;
#pragma unroll
        for (int w = 0; w < 8; w++) {
            const u64 k = 8 * b + w;
            if (k >= draws) break;
            const u64 v = (u64)(x[2 * w] + st[2 * w]) | ((u64)(x[2 * w + 1] + st[2 * w + 1]) << 32);
            if (v * q > zone) bad = true;
            if (k >= keep_from) out[(size_t)s * kept + (k - keep_from)] = __umul64hi(v, q);
        }
        if (bad) reject[s] = 1;
    }
}
#undef PZ_QR

void launch_expand_uniform(const u32 *keys, u32 S, u64 draws, u64 keep_from, u64 q, u64 *out,
                           u32 *reject, hipStream_t s)
{
    const u64 work = (u64)S * ((draws + 7) / 8);
    if (work == 0) return;
    const u64 zone = ~0ull - (0ull - q) % q;  // u64::MAX - (2^64 mod q)
    const u64 g = (work + 255) / 256;
    hipLaunchKernelGGL(k_expand_uniform, (int)(g < 65535 * 8 ? g : 65535 * 8), 256, 0, s, keys,
                       S, draws, keep_from, q, zone, out, reject);
}

}  // namespace pz
