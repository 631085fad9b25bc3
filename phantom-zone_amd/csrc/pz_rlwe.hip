// pz_rlwe.hip -- RLWE operators and the persistent blind rotation (gfx950).
//
//   k_ext             batched rlwe_by_rgsw_shoup
//   k_auto            batched rlwe_auto_shoup
//   k_blind_rotate    persistent: dlog buckets, test-vector rotation, LMKCY+
//                     blind rotation, sample extract (one PBS per work group
//                     at a time, RLWE resident in VGPRs)
#include <algorithm>

#include "pz_internal.hpp"

#ifndef PZ_BR_WAVES
#define PZ_BR_WAVES 2
#endif
#ifndef PZ_BR_SPLIT
#define PZ_BR_SPLIT 1
#endif
// waves per SIMD of the split kernel's two variants: latency (a bootstrap per
// CU) and throughput (two per CU, 128 VGPRs)
#ifndef PZ_BR_SPLIT_WAVES_LAT
#define PZ_BR_SPLIT_WAVES_LAT 2
#endif
#ifndef PZ_BR_SPLIT_WAVES
#define PZ_BR_SPLIT_WAVES 4
#endif
// the same for the split kernel with 4 coefficients per thread (twice the
// waves per bootstrap)
#ifndef PZ_BR_SPLIT4_WAVES_LAT
#define PZ_BR_SPLIT4_WAVES_LAT 4
#endif
#ifndef PZ_BR_SPLIT4_WAVES
#define PZ_BR_SPLIT4_WAVES 8
#endif
#define PZ_BR_ATTR __attribute__((amdgpu_waves_per_eu(PZ_BR_WAVES, PZ_BR_WAVES)))

namespace pz {

// ---------------------------------------------------------------------------
// Batched RLWE x RGSW / automorphism (one RLWE per work group).
// rlwe: [B][2][N]; idx[b] selects the RGSW (LWE index) or auto key.
// ---------------------------------------------------------------------------
template <int LOGN>
__global__ __launch_bounds__(1 << (LOGN - 3)) PZ_BR_ATTR void k_ext(u64 *rlwe, const u32 *idx,
                                                         const u8 *trivial, size_t B, DevCtx c,
                                                         DevKey key)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ u64 lds[];
    const u32 j = threadIdx.x;
    for (size_t bb = blockIdx.x; bb < B; bb += gridDim.x) {
        u64 *p = rlwe + bb * 2 * N;
        u64 a[8], b[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            a[r] = p[coef_x<LOGN>(j, r)];
            b[r] = p[N + coef_x<LOGN>(j, r)];
        }
        if (idx[bb] >= c.n_lwe) continue;  // no such key: unchanged
        size_t off = (size_t)idx[bb] * key.rgsw_rows * N;
        ext_product_blk<LOGN>(a, b, key.rgsw + off, key.rgsw_s + off, trivial[bb] != 0, c, j, lds);
#pragma unroll
        for (int r = 0; r < 8; r++) {
            p[coef_x<LOGN>(j, r)] = a[r];
            p[N + coef_x<LOGN>(j, r)] = b[r];
        }
    }
}

template <int LOGN>
__global__ __launch_bounds__(1 << (LOGN - 3)) void k_auto(u64 *rlwe, const u32 *kidx,
                                                          const u8 *trivial, size_t B, DevCtx c,
                                                          DevKey key)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ u64 lds[];
    const u32 j = threadIdx.x;
    for (size_t bb = blockIdx.x; bb < B; bb += gridDim.x) {
        u64 *p = rlwe + bb * 2 * N;
        u64 a[8], b[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            a[r] = p[coef_x<LOGN>(j, r)];
            b[r] = p[N + coef_x<LOGN>(j, r)];
        }
        u32 k = kidx[bb];
        if (k > c.w) continue;  // no such key: unchanged
        size_t off = (size_t)k * key.auto_rows * N;
        auto_blk<LOGN>(a, b, key.auto_k + off, key.auto_s + off, c.auto_el[k], trivial[bb] != 0,
                       c, j, lds);
#pragma unroll
        for (int r = 0; r < 8; r++) {
            p[coef_x<LOGN>(j, r)] = a[r];
            p[N + coef_x<LOGN>(j, r)] = b[r];
        }
    }
}

// ---------------------------------------------------------------------------
// Order of a batch by key index: a counting sort in one work group (counts,
// exclusive prefix sum and placement in LDS; indices past the last key share
// the final counter).  The order within one index is arbitrary, and each
// RLWE's result does not depend on it.  nkeys + 1 <= KEY_ORDER_MAX.
// ---------------------------------------------------------------------------
constexpr u32 KEY_ORDER_MAX = 8192;
__global__ __launch_bounds__(1024) void k_key_order(const u32 *idx, size_t B, u32 nkeys,
                                                    u32 *order)
{
    __shared__ u32 cnt[KEY_ORDER_MAX];
    __shared__ u32 wsum[16];
    const u32 t = threadIdx.x, n = nkeys + 1, lane = t & 63, wv = t >> 6;
    for (u32 i = t; i < n; i += 1024) cnt[i] = 0;
    __syncthreads();
    for (size_t b = t; b < B; b += 1024) atomicAdd(&cnt[idx[b] < nkeys ? idx[b] : nkeys], 1u);
    __syncthreads();
    // exclusive prefix sum: thread t takes ceil(n / 1024) consecutive
    // counters; their sums are scanned within each wave (shuffles), then the
    // 16 wave totals by wave 0
    const u32 per = (n + 1023) / 1024, lo = t * per;
    u32 sum = 0;
    for (u32 i = 0; i < per && lo + i < n; i++) sum += cnt[lo + i];
    u32 v = sum;
#pragma unroll
    for (u32 off = 1; off < 64; off <<= 1) {
        const u32 y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    if (wv == 0) {
        u32 w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (u32 off = 1; off < 16; off <<= 1) {
            const u32 y = __shfl_up(w, off, 64);
            if (lane >= off) w += y;
        }
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    u32 run = v - sum + (wv ? wsum[wv - 1] : 0);
    for (u32 i = 0; i < per && lo + i < n; i++) {
        const u32 c = cnt[lo + i];
        cnt[lo + i] = run;
        run += c;
    }
    __syncthreads();
    for (size_t b = t; b < B; b += 1024)
        order[atomicAdd(&cnt[idx[b] < nkeys ? idx[b] : nkeys], 1u)] = (u32)b;
}

bool launch_key_order(const u32 *idx, size_t B, u32 nkeys, u32 *order, hipStream_t s)
{
    if (nkeys + 1 > KEY_ORDER_MAX) return false;
    hipLaunchKernelGGL(k_key_order, 1, 1024, 0, s, idx, B, nkeys, order);
    return true;
}

// ---------------------------------------------------------------------------
// Batched RLWE x RGSW / automorphism on the split work group (d = 1 sets):
// the blind rotation's step, one RLWE per 2 x N/8 threads; half h loads and
// stores polynomial h.  LDS: the two halves' regions and the stash.
// ---------------------------------------------------------------------------
// Persistent.  With `order` (the batch sorted by key index) the positions are
// dealt XCD by XCD: work groups are dispatched to the 8 XCDs in turn (work
// group g on XCD g % 8), so round t gives the G/8 work groups of XCD x the
// consecutive positions [(8t + x) G/8, (8t + x + 1) G/8) -- a few key
// indices per XCD at a time, whose key slices the XCD's L2 then serves to
// every RLWE of that index.  Without an order, contiguous runs.
template <int LOGN>
__global__ __launch_bounds__(2 << (LOGN - 3)) __attribute__((amdgpu_waves_per_eu(
    PZ_BR_SPLIT_WAVES, PZ_BR_SPLIT_WAVES))) void k_step_split(bool is_auto, u64 *rlwe,
                                                                const u32 *idx, const u8 *trivial,
                                                                const u32 *order, size_t B,
                                                                DevCtx c, DevKey key)
{
    constexpr int N = 1 << LOGN;
    constexpr int TPB = N / 8;
    constexpr int W = Geom<LOGN>::LDS_WORDS;
    extern __shared__ u64 lds[];
    const u32 tid = threadIdx.x, j = tid % TPB;
    const u32 h = (u32)__builtin_amdgcn_readfirstlane(tid / TPB);
    const size_t rs = (size_t)(is_auto ? key.auto_rows : key.rgsw_rows) * N;
    constexpr u32 XCDS = 8;
    const bool dealt = order && gridDim.x % XCDS == 0;
    const size_t per_x = gridDim.x / XCDS, x = blockIdx.x % XCDS, m = blockIdx.x / XCDS;
    const size_t p0 = dealt ? 0 : blockIdx.x * B / gridDim.x;
    const size_t p1 = dealt ? B : (blockIdx.x + 1) * B / gridDim.x;
    // A result waits in this thread's own words of its half's LDS region
    // (the stride-layout words its inverse transform loaded last, which the
    // next forward transform stores to first) and is written out after the
    // next RLWE's loads are issued: loads and stores share one in-order
    // counter, so a load issued after the stores would wait for them.
    u64 *mine = lds + h * W;
    u64 *parked = nullptr;
    for (size_t it = p0; it < p1; it++) {
        size_t pos = it;
        if (dealt) {
            pos = ((it * XCDS) + x) * per_x + m;  // round it of this work group
            if (pos >= B) break;
        }
        const size_t bb = order ? (size_t)__builtin_amdgcn_readfirstlane((int)order[pos]) : pos;
        const u32 k = (u32)__builtin_amdgcn_readfirstlane((int)idx[bb]);
        if (k >= (is_auto ? c.w + 1 : c.n_lwe)) continue;  // no such key: unchanged
        u64 *poly = rlwe + (bb * 2 + h) * N;
        u64 p[8];
#pragma unroll
        for (int r = 0; r < 8; r++) p[r] = poly[coef_x<LOGN>(j, r)];
        if (parked) {
#pragma unroll
            for (int r = 0; r < 8; r++)
                parked[coef_x<LOGN>(j, r)] = mine[lds_ix(coef_x<LOGN>(j, r))];
        }
        const bool triv = __builtin_amdgcn_readfirstlane((int)trivial[bb]) != 0;
        step_split<LOGN, 8>(p, (is_auto ? key.auto_k : key.rgsw) + k * rs,
                            (is_auto ? key.auto_s : key.rgsw_s) + k * rs, is_auto,
                            c.auto_el[is_auto ? k : 0], triv, c, j, h, lds, lds + 2 * W);
#pragma unroll
        for (int r = 0; r < 8; r++) mine[lds_ix(coef_x<LOGN>(j, r))] = p[r];
        parked = poly;
    }
    if (parked) {
#pragma unroll
        for (int r = 0; r < 8; r++) parked[coef_x<LOGN>(j, r)] = mine[lds_ix(coef_x<LOGN>(j, r))];
    }
}

// ---------------------------------------------------------------------------
// RGSW x RGSW (rgsw_by_rgsw_inplace, src/rgsw/runtime.rs:883-985): each work
// group takes one RLWE row pair of one RGSW0 (coefficient domain; rows (i,
// da0 + i) of the -sm part, then (2 da0 + i, 2 da0 + db0 + i) of the m part)
// and multiplies it by that RGSW0's RGSW1 (evaluation domain, rows scaled by
// n^-1, Shoup twins) with the RGSW x RGSW decomposers in c.dec_a / c.dec_b.
// ---------------------------------------------------------------------------
template <int LOGN>
__global__ __launch_bounds__(1 << (LOGN - 3)) PZ_BR_ATTR void k_rgsw_rows(
    u64 *rgsw0, u32 r0, u32 da0, u32 db0, size_t B, DevCtx c, const u64 *K1, const u64 *K1s,
    u32 r1)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ u64 lds[];
    const u32 j = threadIdx.x;
    const u32 npairs = da0 + db0;
    for (size_t w = blockIdx.x; w < B * npairs; w += gridDim.x) {
        const size_t bb = w / npairs;
        const u32 k = (u32)(w % npairs);
        const u32 ra = k < da0 ? k : 2 * da0 + (k - da0);
        const u32 rb = k < da0 ? da0 + k : 2 * da0 + db0 + (k - da0);
        u64 *pa = rgsw0 + (bb * r0 + ra) * N, *pb = rgsw0 + (bb * r0 + rb) * N;
        u64 a[8], b[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            a[r] = pa[coef_x<LOGN>(j, r)];
            b[r] = pb[coef_x<LOGN>(j, r)];
        }
        const size_t off = bb * r1 * N;
        ext_product_blk<LOGN>(a, b, K1 + off, K1s + off, false, c, j, lds);
#pragma unroll
        for (int r = 0; r < 8; r++) {
            pa[coef_x<LOGN>(j, r)] = a[r];
            pb[coef_x<LOGN>(j, r)] = b[r];
        }
    }
}

void launch_rgsw_rows(int logn, u64 *rgsw0, u32 r0, u32 da0, u32 db0, size_t B, const DevCtx &c,
                      const u64 *K1, const u64 *K1s, u32 r1, hipStream_t s)
{
    const size_t work = B * (size_t)(da0 + db0);
    if (work == 0) return;
    const int grid = (int)(work < 65535 ? work : 65535);
    const size_t lds = 2 * 8 * (size_t)(((size_t)1 << logn) + ((size_t)1 << logn) / 8);
    switch (logn) {
    case 9:
        hipLaunchKernelGGL(k_rgsw_rows<9>, grid, 1 << 6, lds, s, rgsw0, r0, da0, db0, B, c, K1,
                           K1s, r1);
        break;
    case 11:
        hipLaunchKernelGGL(k_rgsw_rows<11>, grid, 1 << 8, lds, s, rgsw0, r0, da0, db0, B, c, K1,
                           K1s, r1);
        break;
    }
}

// ---------------------------------------------------------------------------
// The blind rotation's op list (src/pbs.rs:240-414), built by the whole work
// group.  The reference walks the dlog buckets k = nb-1 .. 0 (its -(g^k)
// buckets, then its +(g^k) ones: k >= br_q/4 is the first half) and emits each
// bucket's LWE indices in ascending order; within a half, a pending
// automorphism fires after bucket k (as SCHED_AUTO | v, v the buckets walked
// since the last one) when bucket k-1 is non-empty, when v reaches w, or at
// the half's bucket 1; the first half ends with its bucket br_q/4 and the -g
// automorphism (v = 0).  Whether one fires after bucket k is a function of the
// bucket sizes alone: with d the distance from k to the half's last natural
// firing point above it (one past the next non-empty bucket at or above k,
// below the half's top bucket; the half's top + 1 if none), v = (d-1) mod w + 1
// and a forced firing happens where d mod w = 0.  Each bucket's op count is
// then known, and one exclusive scan over them places every op.
// ---------------------------------------------------------------------------

// 0 when no automorphism follows bucket k, else 1 + v
PZ_DEV u32 br_auto_after(u32 k, u32 qb4, u32 w, const u32 *cnt, const u64 *occ)
{
    const u32 base = k >= qb4 ? qb4 : 0, i = k - base;
    if (i == 0) return base ? 1u : 0u;
    const bool nat = i == 1 || cnt[k - 1] != 0;
    const u32 top = base + qb4 - 1;  // non-empty buckets searched in [k, top)
    u32 d = top + 1 - k;
    u32 kk = k;
    u64 m = occ[k / 64] >> (k % 64);
    for (;;) {
        if (m) {
            const u32 nz = kk + (u32)__builtin_ctzll(m);
            if (nz < top) d = nz + 1 - k;
            break;
        }
        kk = (kk / 64 + 1) * 64;
        if (kk >= top) break;
        m = occ[kk / 64];
    }
    return nat || d % w == 0 ? 2u + (d - 1) % w : 0u;
}

// (3) odd moddown of the key-switched LWE to br_q, dlog buckets, and (5) the
// op list into sched[0 .. *nsched).  Scratch in the polynomial area of LDS:
// bucket sizes, non-empty bitmap, per-wave sums, bucket and rank of each LWE
// index.  bpos: [nb] first op of each bucket.
template <int THREADS>
PZ_DEV void br_schedule(const u32 *in, const DevCtx &c, u32 tid, u64 *lds, u16 *bpos,
                        u16 *sched, u32 *nsched)
{
    constexpr u32 WAVES = THREADS / 64;
    const u32 nb = c.br_q >> 1, qb4 = c.br_q >> 2, n = c.n_lwe, w = c.w;
    u32 *cnt = reinterpret_cast<u32 *>(lds);            // [nb]
    u64 *occ = lds + nb / 2;                             // [nb / 64]
    u32 *wsum = reinterpret_cast<u32 *>(occ + nb / 64);  // [WAVES]
    u16 *kbin = reinterpret_cast<u16 *>(wsum + WAVES);   // [n]
    u16 *rk = kbin + n;                                  // [n]
    for (u32 k = tid; k < nb; k += THREADS) cnt[k] = 0;
    __syncthreads();
    for (u32 i = tid; i < n; i += THREADS) {
        const u32 v = in[1 + i] & (u32)c.lwe_mask;
        const u32 odd = ((u32)(((u64)v * c.br_q) >> c.lwe_q_log)) | 1u;
        const u32 k = c.dlog[odd];
        kbin[i] = (u16)k;
        rk[i] = (u16)atomicAdd(&cnt[k], 1u);  // arrival order, ranked below
    }
    __syncthreads();
    // non-empty buckets, 64 per wave (nb is a multiple of 64)
    for (u32 k0 = tid & ~63u; k0 < nb; k0 += THREADS) {
        const u64 m = __ballot(cnt[k0 + (tid & 63)] != 0);
        if ((tid & 63) == 0) occ[k0 / 64] = m;
    }
    __syncthreads();
    // thread t takes the E buckets nb-1-t*E .. nb-(t+1)*E (op-list order)
    const u32 E = nb > THREADS ? nb / THREADS : 1u;
    const u32 p0 = tid * E;
    u32 s = 0;
    for (u32 e = 0; e < E; e++) {
        const u32 p = p0 + e;
        if (p < nb) s += cnt[nb - 1 - p] + (br_auto_after(nb - 1 - p, qb4, w, cnt, occ) != 0);
    }
    const u32 lane = tid & 63, wv = tid >> 6;
    u32 x = s;  // inclusive scan over the wave
#pragma unroll
    for (u32 o = 1; o < 64; o <<= 1) {
        const u32 y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    u32 pos = x - s;
    for (u32 v = 0; v < wv; v++) pos += wsum[v];
    if (tid == THREADS - 1) *nsched = pos + s;
    for (u32 e = 0; e < E; e++) {
        const u32 p = p0 + e;
        if (p < nb) {
            const u32 k = nb - 1 - p;
            const u32 a = br_auto_after(k, qb4, w, cnt, occ);
            bpos[k] = (u16)pos;
            pos += cnt[k];
            if (a) sched[pos++] = (u16)(SCHED_AUTO | (a - 1));
        }
    }
    __syncthreads();
    // each bucket's indices in arrival order, then each index to its rank
    // among them (ascending, the reference's order)
    for (u32 i = tid; i < n; i += THREADS) sched[bpos[kbin[i]] + rk[i]] = (u16)i;
    __syncthreads();
    for (u32 i = tid; i < n; i += THREADS) {
        const u32 k = kbin[i], b0 = bpos[k], nk = cnt[k];
        u32 r = 0;
        for (u32 t = 0; t < nk; t++) r += sched[b0 + t] < i;
        rk[i] = (u16)r;
    }
    __syncthreads();
    for (u32 i = tid; i < n; i += THREADS) sched[bpos[kbin[i]] + rk[i]] = (u16)i;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// k_blind_rotate: persistent PBS tail (src/pbs.rs:146-232).
//   ks:   [B][n+1] key-switched LWE (mod q_ks)
//   gate: [B] test-vector index
//   out:  [B][N+1] output LWE (mod Q)
// VAR 0 (BR_ONE): N/8 threads hold a and b (16 coefficients each, any d).
// VAR 1, 2 (BR_SPLIT_LAT, BR_SPLIT): 2 x N/8 threads, half h holds
// polynomial h (d = 1 only); register budgets for 2 and 4 waves per SIMD.
// VAR 3, 4 (BR_SPLIT4_LAT, BR_SPLIT4): 2 x N/4 threads, 4 coefficients per
// thread (radix-4 register phases); budgets for 4 and 8 waves per SIMD.
// LDS: 2 polynomials (exchange / permutation; SPLIT: one per half), SPLIT: a
// third for the permuted b of an automorphism; bucket positions + op list.
// ---------------------------------------------------------------------------
template <int VAR>
constexpr int br_waves()
{
    return VAR == BR_SPLIT ? PZ_BR_SPLIT_WAVES
         : VAR == BR_SPLIT_LAT ? PZ_BR_SPLIT_WAVES_LAT
         : VAR == BR_SPLIT4 ? PZ_BR_SPLIT4_WAVES
         : VAR == BR_SPLIT4_LAT ? PZ_BR_SPLIT4_WAVES_LAT : PZ_BR_WAVES;
}
template <int VAR>
constexpr int br_coefs() { return VAR == BR_SPLIT4 || VAR == BR_SPLIT4_LAT ? 4 : 8; }

template <int LOGN, int VAR>
__global__ __launch_bounds__((VAR ? 2 : 1) * (1 << LOGN) / br_coefs<VAR>())
    __attribute__((amdgpu_waves_per_eu(br_waves<VAR>(), br_waves<VAR>()))) void
k_blind_rotate(
    const u32 *ks, const u32 *gate, size_t B, u64 *out, u32 *stats, DevCtx c, DevKey key0,
    const DevKey *keys, const u32 *key_idx, const u32 *out_idx)
{
    constexpr bool SPLIT = VAR != BR_ONE;
    constexpr int R = br_coefs<VAR>();  // coefficients per thread and polynomial
    constexpr int N = 1 << LOGN;
    constexpr int TPB = N / R;
    constexpr int THREADS = SPLIT ? 2 * TPB : TPB;
    constexpr int W = Geom<LOGN>::LDS_WORDS;
    extern __shared__ u64 lds[];
    const u32 nb = c.br_q >> 1;
    const u32 n = c.n_lwe;
    u64 *stash = lds + 2 * W;                               // SPLIT: permuted b
    u32 *nsched = reinterpret_cast<u32 *>(lds + (SPLIT ? 3 : 2) * W);  // op count
    u16 *bpos = reinterpret_cast<u16 *>(nsched + 2);        // [nb] first op of each bucket
    u16 *sched = bpos + nb;                                 // [br_sched_len] op list
    const u32 tid = threadIdx.x;
    const u32 j = SPLIT ? tid % TPB : tid;
    // half index, uniform over each wave (TPB is a multiple of 64)
    const u32 h = SPLIT ? (u32)__builtin_amdgcn_readfirstlane(tid / TPB) : 0u;
    const u64 q = c.tb.q;

    for (size_t bb = blockIdx.x; bb < B; bb += gridDim.x) {
        const u32 *in = ks + bb * (n + 1);
        // the bootstrap's server key: one key for the batch, or one per
        // ciphertext (sessions batched together)
        const DevKey key = keys ? uniform_key(keys[key_idx[bb]]) : key0;
        br_schedule<THREADS>(in, c, tid, lds, bpos, sched, nsched);

        // (4) test vector times X^{g*b} (monomial_mul, src/pbs.rs:163-211)
        const u32 oddb = ((u32)(((u64)(in[0] & (u32)c.lwe_mask) * c.br_q) >> c.lwe_q_log)) | 1u;
        u32 exp = (u32)(((u64)c.g * oddb) % c.br_q);
        bool msign = true;
        if (exp > nb) {
            exp -= nb;
            msign = false;
        }
        // gate ids outside 0..5 are invalid input; clamp so memory stays in bounds
        const u32 gid = gate[bb] < 6 ? gate[bb] : 0;
        const u64 *tv = c.test_vec + (size_t)gid * nb;
        u64 tvr[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            u32 x = coef_x<LOGN, R>(j, r);
            u64 val = 0;
            if (x % c.embedding == 0) {
                u32 y = x / c.embedding;
                u32 src;
                bool sg;
                if (y >= exp) {
                    src = y - exp;
                    sg = msign;
                } else {
                    src = y + nb - exp;
                    sg = !msign;
                }
                val = sg ? tv[src] : neg_mod(tv[src], q);
            }
            tvr[r] = val;
        }
        // accumulator RLWE (0, test vector): split halves hold a or b in p
        u64 a[R], b[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            a[r] = SPLIT ? (h ? tvr[r] : 0) : 0;  // SPLIT: a[] is this half's polynomial
            b[r] = tvr[r];
        }

        // (5) blind rotation: the op list built above, one step at a time
        const size_t rs = (size_t)key.rgsw_rows * N, as = (size_t)key.auto_rows * N;
        bool trivial = true;
        u32 n_auto = 0;
        const u32 nops = (u32)__builtin_amdgcn_readfirstlane((int)*nsched);
        for (u32 k = 0; k < nops; k++) {
            const u32 op = (u32)__builtin_amdgcn_readfirstlane((int)sched[k]);
            const bool is_auto = (op & SCHED_AUTO) != 0;
            const u32 arg = op & (SCHED_AUTO - 1);
            if constexpr (SPLIT) {
                const size_t off = is_auto ? arg * as : arg * rs;
                step_split<LOGN, R, false>(a, (is_auto ? key.auto_k : key.rgsw) + off,
                                 (is_auto ? key.auto_s : key.rgsw_s) + off, is_auto,
                                 c.auto_el[is_auto ? arg : 0], trivial, c, j, h, lds, stash);
            } else if constexpr (R == 8) {
                if (is_auto)
                    auto_blk<LOGN>(a, b, key.auto_k + arg * as, key.auto_s + arg * as,
                                   c.auto_el[arg], trivial, c, j, lds);
                else
                    ext_product_blk<LOGN>(a, b, key.rgsw + arg * rs, key.rgsw_s + arg * rs,
                                          trivial, c, j, lds);
            }
            n_auto += is_auto;
            trivial = trivial && is_auto;
        }

        // (6) sample_extract at index 0 (src/pbs.rs:423-449); the split
        // steps leave the RLWE in [0, 2q) (PZ_LOOSE_STEPS): reduced here
        if constexpr (SPLIT && PZ_LOOSE_STEPS) {
#pragma unroll
            for (int r = 0; r < R; r++) a[r] = umin(a[r], a[r] - q);
        }
        if (stats && tid == 0) stats[bb] = n_auto;
        u64 *o = out + (out_idx ? (size_t)out_idx[bb] : bb) * (size_t)(N + 1);
        const bool has_a = !SPLIT || h == 0, has_b = !SPLIT || h == 1;
#pragma unroll
        for (int r = 0; r < R; r++) {
            u32 x = coef_x<LOGN, R>(j, r);
            const u64 av = a[r], bv = SPLIT ? a[r] : b[r];
            if (x == 0) {
                if (has_a) o[1] = av;
                if (has_b) o[0] = bv;
            } else if (has_a) {
                o[1 + N - x] = neg_mod(av, q);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <int LOGN>
static void launch_ext_t(bool is_auto, u64 *rlwe, const u32 *idx, const u8 *triv, size_t B,
                         const DevCtx &c, const DevKey &k, hipStream_t s)
{
    constexpr int TPB = 1 << (LOGN - 3);
    size_t lds = 2 * 8 * (size_t)Geom<LOGN>::LDS_WORDS;
    int grid = B < 65535 ? (int)B : 65535;
    if (grid < 1) return;
    if (is_auto)
        hipLaunchKernelGGL((k_auto<LOGN>), grid, TPB, lds, s, rlwe, idx, triv, B, c, k);
    else
        hipLaunchKernelGGL((k_ext<LOGN>), grid, TPB, lds, s, rlwe, idx, triv, B, c, k);
}

// Resident work groups of the split step per CU (occupancy query, cached).
int step_split_per_cu(int logn)
{
    static int per[2] = {0, 0};
    int &v = per[logn == 11 ? 1 : 0];
    if (v == 0) {
        const size_t lds = 3 * 8 * (size_t)((1 << logn) + (1 << logn) / 8);
        int n = 0;
        const hipError_t e =
            logn == 11 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step_split<11>, 2 << 8, lds)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step_split<9>, 2 << 6, lds);
        v = (e == hipSuccess && n > 0) ? n : 1;
    }
    return v;
}

void launch_ext(int logn, bool is_auto, u64 *rlwe, const u32 *idx, const u8 *triv, size_t B,
                const DevCtx &c, const DevKey &k, hipStream_t s, const u32 *order, int grid)
{
    if (br_split(c)) {  // d = 1: the blind rotation's split step, persistent
        if (B == 0) return;
        if (grid <= 0 || (size_t)grid > B) grid = B < 65535 ? (int)B : 65535;
        const size_t lds = 3 * 8 * (size_t)((1 << logn) + (1 << logn) / 8);
        if (logn == 11)
            hipLaunchKernelGGL(k_step_split<11>, grid, 2 << 8, lds, s, is_auto, rlwe, idx, triv,
                               order, B, c, k);
        else if (logn == 9)
            hipLaunchKernelGGL(k_step_split<9>, grid, 2 << 6, lds, s, is_auto, rlwe, idx, triv,
                               order, B, c, k);
        return;
    }
    switch (logn) {
    case 9: launch_ext_t<9>(is_auto, rlwe, idx, triv, B, c, k, s); break;
    case 11: launch_ext_t<11>(is_auto, rlwe, idx, triv, B, c, k, s); break;
    }
}

size_t br_lds_bytes(int logn, bool split, u32 br_q, u32 n_lwe)
{
    size_t words = (size_t)(1 << logn) + (size_t)(1 << logn) / 8;
    size_t poly = (split ? 3 : 2) * 8 * words;
    // br_schedule's scratch: bucket sizes, non-empty bitmap, wave sums (at
    // most 16 waves), bucket and rank of each LWE index
    size_t scratch_need = 4 * (size_t)(br_q / 2) + br_q / 16 + 64 + 4 * (size_t)n_lwe;
    if (scratch_need > poly) poly = scratch_need;
    // op count (8 bytes), bucket positions, op list (u16 each)
    return poly + 8 + 2 * ((size_t)(br_q / 2) + br_sched_len(br_q, n_lwe)) + 16;
}

// The split work group serves d_a = d_b = d_auto = 1 (every NI set).
bool br_split(const DevCtx &c) { return PZ_BR_SPLIT && c.da == 1 && c.db == 1 && c.dauto == 1; }

int br_threads(int logn, int var)
{
    return (var ? 2 : 1) << (logn - (var == BR_SPLIT4 || var == BR_SPLIT4_LAT ? 2 : 3));
}

template <int LOGN, int VAR>
static const void *br_kernel_ptr()
{
    return reinterpret_cast<const void *>(&k_blind_rotate<LOGN, VAR>);
}

template <int LOGN>
static const void *br_kernel_of(int var)
{
    switch (var) {
    case BR_SPLIT: return br_kernel_ptr<LOGN, BR_SPLIT>();
    case BR_SPLIT_LAT: return br_kernel_ptr<LOGN, BR_SPLIT_LAT>();
    case BR_SPLIT4: return br_kernel_ptr<LOGN, BR_SPLIT4>();
    case BR_SPLIT4_LAT: return br_kernel_ptr<LOGN, BR_SPLIT4_LAT>();
    default: return br_kernel_ptr<LOGN, BR_ONE>();
    }
}

const void *blind_rotate_kernel(int logn, int var)
{
    switch (logn) {
    case 9: return br_kernel_of<9>(var);
    case 11: return br_kernel_of<11>(var);
    }
    return nullptr;
}

void launch_blind_rotate(int logn, int var, int grid, const u32 *ks, const u32 *gate, size_t B,
                         u64 *out, u32 *stats, const DevCtx &c, const DevKey &k,
                         const DevKey *keys, const u32 *key_idx, const u32 *out_idx,
                         hipStream_t s)
{
    if (B == 0) return;
    if (!br_split(c)) var = BR_ONE;
    size_t lds = br_lds_bytes(logn, var != BR_ONE, c.br_q, c.n_lwe);
    const int tpb = br_threads(logn, var);
#define PZ_BR_LAUNCH(L, V)                                                                       \
    hipLaunchKernelGGL((k_blind_rotate<L, V>), grid, tpb, lds, s, ks, gate, B, out, stats, c, k, \
                       keys, key_idx, out_idx)
#define PZ_BR_LAUNCH_ALL(L)                                                                      \
    switch (var) {                                                                               \
    case BR_SPLIT: PZ_BR_LAUNCH(L, BR_SPLIT); break;                                             \
    case BR_SPLIT_LAT: PZ_BR_LAUNCH(L, BR_SPLIT_LAT); break;                                     \
    case BR_SPLIT4: PZ_BR_LAUNCH(L, BR_SPLIT4); break;                                           \
    case BR_SPLIT4_LAT: PZ_BR_LAUNCH(L, BR_SPLIT4_LAT); break;                                   \
    default: PZ_BR_LAUNCH(L, BR_ONE); break;                                                     \
    }
    switch (logn) {
    case 9: PZ_BR_LAUNCH_ALL(9); break;
    case 11: PZ_BR_LAUNCH_ALL(11); break;
    }
#undef PZ_BR_LAUNCH_ALL
#undef PZ_BR_LAUNCH
}

}  // namespace pz
