// pz_ni.hip -- server-side aggregation of non-interactive multi-party
// server-key shares on the device.
//
//   pz_ni_uitos_ksk       Ksk(u_j -> s): A rows expanded from user j's seed,
//                         B = ksk_b + the others' zero encryptions, both sent
//                         to the evaluation domain (src/bool/evaluator.rs:
//                         1266-1337)
//   k_ni_rows             produce_rgsw_ciphertext_from_ni_rgsw
//                         (src/bool/evaluator.rs:480-608): one work group per
//                         produced RLWE (LWE index, party, row).  The input
//                         polynomial (sum of the parties' zero encryptions
//                         for RLWE'(-sm) rows, -a_{i,l} for RLWE'(m) rows) is
//                         decomposed in registers under the u -> s decomposer;
//                         each digit row is transformed and multiply-
//                         accumulated with the party's Ksk rows.  The index
//                         leader's rows go back to the coefficient domain
//                         (plus its NI row); the others' stay in the
//                         evaluation domain (plus the transformed NI row),
//                         scaled by n^-1 for the RGSW x RGSW step
//   pz_ni_aggregate_rgsw  k_ni_rows, Shoup twins of the others' RGSWs, then
//                         one RGSW x RGSW launch per non-leader slot, in
//                         ascending party order (src/bool/evaluator.rs:
//                         1511-1560)
#include <vector>

#include "pz_host.hpp"
#include "pz_internal.hpp"

#ifndef PZ_NI_GROUP
#define PZ_NI_GROUP 3
#endif

namespace pz {

struct NiArgs {
    const u64 *zero;   // [P][count][da][N]
    const u64 *cts;    // [P][count][dmax][N]
    const u64 *arows;  // [count][db][N]: a_{i,l}, i in [dmax - db, dmax)
    const u64 *K[PZ_NI_MAX_PARTIES], *Ks[PZ_NI_MAX_PARTIES];
    u64 *out0;         // [count][2 da0 + 2 db0][N], coefficient domain
    u64 *out1;         // [P - 1][count][2 da + 2 db][N], evaluation domain * n^-1
    u32 P, count, l0, per, da, db, dmax, da0, db0, d;
};

// Consumers of RGSW x RGSW row r of an index: the non-leaders (each produces
// every row), then the leader when r is one of its rows (its counts da0 <= da,
// db0 <= db take the last rows of each part).
PZ_DEV u32 ni_consumers(const NiArgs &A, u32 r)
{
    const bool lead_uses = r < A.da ? r >= A.da - A.da0 : r - A.da >= A.db - A.db0;
    return A.P - 1 + (lead_uses ? 1 : 0);
}

// One work group per (index, input polynomial, group of up to G consumers):
// the 50 digit rows of the polynomial are transformed once for the whole
// group and multiply-accumulated with each member's Ksk.  With PZ_NI_LAZY the
// products are [0,4q) Shoup products left unreduced (the accumulator stays
// below 4q times the row count, < 2^62 for d <= 64) and one Shoup multiply by
// 1 brings each accumulator to [0,2q) at the end: the same residues, and the
// strict outputs are the same words.
#ifndef PZ_NI_LAZY
#define PZ_NI_LAZY 1
#endif
PZ_DEV void ni_mac_row(u64 (&acc)[8], const u64 *__restrict__ K, const u64 *__restrict__ Ks,
                       const u64 (&D)[8], u32 j, u64 q, u64 q2)
{
#if PZ_NI_LAZY
    fma_row_lazy4(acc, K, Ks, D, j, q, q2);
#else
    fma_row<false>(acc, K, Ks, D, j, q, q2);
#endif
}
// accumulator of ni_mac_row in [0,2q)
PZ_DEV u64 ni_mac_done(u64 a, const Tables &tb)
{
#if PZ_NI_LAZY
    return fma_done(a, tb);
#else
    return a;
#endif
}
template <int LOGN, int G>
__global__ __launch_bounds__(1 << (LOGN - 3)) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_ni_rows(NiArgs A, DevCtx c)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ u64 lds[];
    const u32 j = threadIdx.x;
    const u64 q = c.tb.q, q2 = c.tb.q2;
    const u32 n0 = A.da0 + A.db0, n1 = A.da + A.db;
    u32 per_t = 0;
    for (u32 r = 0; r < n1; r++) per_t += (ni_consumers(A, r) + G - 1) / G;
    for (size_t w = blockIdx.x; w < (size_t)A.count * per_t; w += gridDim.x) {
        const u32 t = (u32)(w / per_t);
        u32 rem = (u32)(w % per_t), r = 0, cons = 0;
        for (;; r++) {
            cons = ni_consumers(A, r);
            const u32 ng = (cons + G - 1) / G;
            if (rem < ng) break;
            rem -= ng;
        }
        const u32 first = rem * G;
        const u32 nmem = cons - first < (u32)G ? cons - first : (u32)G;
        const u32 lead = (A.l0 + t) / A.per;
        const bool nsm = r < A.da;
        const u32 i = nsm ? A.dmax - A.da + r : A.dmax - A.db + (r - A.da);
        u64 val[8];
        if (nsm) {
            // sum over parties of a_{i,l} * s_p + e: a_{i,l} * s + e
#pragma unroll
            for (int x = 0; x < 8; x++) val[x] = 0;
            for (u32 p = 0; p < A.P; p++) {
                const u64 *z = A.zero + (((size_t)p * A.count + t) * A.da + r) * N;
#pragma unroll
                for (int x = 0; x < 8; x++) val[x] = add_mod(val[x], z[coef_x<LOGN>(j, x)], q);
            }
        } else {
            const u64 *a = A.arows + ((size_t)t * A.db + (r - A.da)) * N;
#pragma unroll
            for (int x = 0; x < 8; x++) val[x] = neg_mod(a[coef_x<LOGN>(j, x)], q);
        }
#pragma unroll
        for (int x = 0; x < 8; x++) val[x] = decomp_prepare(c.dec_uitos, val[x]);
        u32 party[G];
#pragma unroll
        for (int m = 0; m < G; m++) {
            const u32 ci = first + m;  // consumer index; P - 1 is the leader
            party[m] = ci >= A.P - 1 ? lead : (ci < lead ? ci : ci + 1);
        }
        u64 acc[G][2][8];
#pragma unroll
        for (int m = 0; m < G; m++)
#pragma unroll
            for (int x = 0; x < 8; x++) acc[m][0][x] = acc[m][1][x] = 0;
        for (u32 row = 0; row < A.d; row++) {
            __builtin_amdgcn_sched_barrier(0);  // no key loads hoisted across rows
            u64 D[1][8];
#pragma unroll
            for (int x = 0; x < 8; x++) D[0][x] = decomp_next(c.dec_uitos, val[x]);
            ntt_fwd<LOGN, 1, PZ_NI_LAZY ? FWD_FAST4 : PZ_FWD_MODE, PZ_NI_LAZY != 0>(D, j, lds, c.tb);
#pragma unroll
            for (int m = 0; m < G; m++) {
                if ((u32)m >= nmem) break;
                const u64 *K = A.K[party[m]], *Ks = A.Ks[party[m]];
                ni_mac_row(acc[m][0], K + (size_t)row * N, Ks + (size_t)row * N, D[0], j, q, q2);
                ni_mac_row(acc[m][1], K + (size_t)(A.d + row) * N, Ks + (size_t)(A.d + row) * N,
                           D[0], j, q, q2);
            }
        }
#pragma unroll
        for (int m = 0; m < G; m++)
#pragma unroll
            for (int x = 0; x < 8; x++) {
                acc[m][0][x] = ni_mac_done(acc[m][0][x], c.tb);
                acc[m][1][x] = ni_mac_done(acc[m][1][x], c.tb);
            }
        const int part = nsm ? 0 : 1;  // the NI row joins a for RLWE'(-sm), b for RLWE'(m)
#pragma unroll
        for (int m = 0; m < G; m++) {
            if ((u32)m >= nmem) break;
            const u32 ci = first + m;
            const u64 *ct = A.cts + (((size_t)party[m] * A.count + t) * A.dmax + i) * N;
            if (ci == A.P - 1) {
                // the leader: RLWE x RGSW counts, coefficient domain
                const u32 k0 = nsm ? r - (A.da - A.da0) : (r - A.da) - (A.db - A.db0);
                const u32 ra = nsm ? k0 : 2 * A.da0 + k0;
                const u32 rb = nsm ? A.da0 + k0 : 2 * A.da0 + A.db0 + k0;
                ntt_inv<LOGN, 2, true, true, false>(acc[m], j, lds, c.tb);
                u64 *o = A.out0 + (size_t)t * 2 * n0 * N;
#pragma unroll
                for (int x = 0; x < 8; x++) {
                    acc[m][part][x] = add_mod(acc[m][part][x], ct[coef_x<LOGN>(j, x)], q);
                    o[(size_t)ra * N + coef_x<LOGN>(j, x)] = acc[m][0][x];
                    o[(size_t)rb * N + coef_x<LOGN>(j, x)] = acc[m][1][x];
                }
            } else {
                // non-leader slot ci: RGSW x RGSW counts, evaluation domain * n^-1
                const u32 ra = nsm ? r : 2 * A.da + (r - A.da);
                const u32 rb = nsm ? A.da + r : 2 * A.da + A.db + (r - A.da);
                u64 cv[1][8];
#pragma unroll
                for (int x = 0; x < 8; x++) cv[0][x] = ct[coef_x<LOGN>(j, x)];
                ntt_fwd<LOGN, 1, FWD_STRICT, false>(cv, j, lds, c.tb);
                u64 *o = A.out1 + (((size_t)ci * A.count + t) * 2 * n1) * N;
#pragma unroll
                for (int x = 0; x < 8; x++) {
                    u64 y0 = umin(acc[m][0][x], acc[m][0][x] - q);
                    u64 y1 = umin(acc[m][1][x], acc[m][1][x] - q);
                    if (part == 0) y0 = add_mod(y0, cv[0][x], q);
                    else y1 = add_mod(y1, cv[0][x], q);
                    y0 = shoup_mul(y0, c.tb.n_inv, c.tb.n_inv_s, q);
                    y1 = shoup_mul(y1, c.tb.n_inv, c.tb.n_inv_s, q);
                    o[(size_t)ra * N + eval_x<LOGN>(j, x)] = umin(y0, y0 - q);
                    o[(size_t)rb * N + eval_x<LOGN>(j, x)] = umin(y1, y1 - q);
                }
            }
        }
    }
}

__global__ void k_ni_sum_rows(u64 *b, const u64 *zero, u32 others, size_t words, u64 q)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 v = b[i];
        for (u32 o = 0; o < others; o++) v = add_mod(v, zero[(size_t)o * words + i], q);
        b[i] = v;
    }
}

static int ni_grid(size_t n) { return (int)((n + 255) / 256 < 65535 * 4 ? (n + 255) / 256 : 65535 * 4); }


}  // namespace pz

using namespace pz;

extern int pz_internal_ctx_view(pz_ctx *c, pz_params *p, DevCtx *dc, int *logn, hipStream_t *s);
extern int pz_internal_rlwe_ksk_adopt(pz_ctx *c, u64 *d_rows, u32 d, pz_rlwe_ksk **out);
extern int pz_internal_workspace(pz_ctx *c, size_t bytes, void **out);
extern int pz_internal_key_adopt(pz_ctx *c, u64 *rgsw, u64 *auto_k, const u64 *ksk_rows,
                                 pz_key **out);

// Uniform rows of S seeded streams into d_out ([S][draws - keep_from]) on
// the device; d_kf is workspace for the keys and the reject flags (12*S
// words).  A stream with a draw outside the acceptance zone (probability
// about 2^-38 per draw for the 54-bit modulus) is regenerated sequentially on
// the host, so the rows always equal the sequential streams.
bool pz_internal_expand_uniform(hipStream_t st, const uint8_t *seeds, u32 S, u64 draws,
                                u64 keep_from, u64 q, u64 *d_out, u32 *d_kf)
{
    u32 *d_keys = d_kf, *d_flags = d_kf + 8 * (size_t)S;
    std::vector<u32> flags(S);
    if (hipMemcpyAsync(d_keys, seeds, 32 * (size_t)S, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(d_flags, 0, 4 * (size_t)S, st) != hipSuccess)
        return false;
    launch_expand_uniform(d_keys, S, draws, keep_from, q, d_out, d_flags, st);
    if (hipMemcpyAsync(flags.data(), d_flags, 4 * (size_t)S, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return false;
    const u64 kept = draws - keep_from;
    std::vector<u64> h(kept);
    for (u32 s = 0; s < S; s++) {
        if (!flags[s]) continue;
        host::ChaCha8 r(seeds + 32 * (size_t)s);
        for (u64 k = 0; k < keep_from; k++) r.uniform(q);
        for (u64 k = 0; k < kept; k++) h[k] = r.uniform(q);
        if (pz_copy_sync(d_out + s * kept, h.data(), 8 * kept, hipMemcpyHostToDevice, st) !=
            hipSuccess)
            return false;
    }
    return true;
}

extern "C" int pz_ni_uitos_ksk(pz_ctx *c, const uint8_t *crs, uint32_t total, uint32_t user,
                               const uint64_t *d_ksk_b, const uint64_t *d_zero_for_user,
                               pz_rlwe_ksk **out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!crs || !d_ksk_b || !out || (total > 1 && !d_zero_for_user))
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !crs || !d_ksk_b || !out || (total > 1 && !d_zero_for_user)");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (p.uitos_d == 0) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.uitos_d == 0");
    if (total == 0 || total > PZ_NI_MAX_PARTIES || user >= total)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: total == 0 || total > PZ_NI_MAX_PARTIES || user >= total");
    const u32 N = p.rlwe_n, d = p.uitos_d;
    const size_t words = (size_t)d * N;
    uint8_t sd[32];
    host::ni_seed_user(crs, user, sd);
    void *ws = nullptr;
    if ((rc = pz_internal_workspace(c, 64, &ws))) return rc;
    u64 *rows = nullptr;
    if (hipMalloc(&rows, 16 * words) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
    bool ok = pz_internal_expand_uniform(st, sd, 1, words, 0, p.rlwe_q, rows, (u32 *)ws) &&
              hipMemcpyAsync(rows + words, d_ksk_b, 8 * words, hipMemcpyDeviceToDevice, st) ==
                  hipSuccess;
    if (ok) {
        if (total > 1)
            hipLaunchKernelGGL(k_ni_sum_rows, ni_grid(words), 256, 0, st, rows + words,
                               d_zero_for_user, total - 1, words, p.rlwe_q);
        launch_ntt(logn, 1, rows, 2 * (size_t)d, dc.tb, st);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    }
    if (ok) rc = pz_internal_rlwe_ksk_adopt(c, rows, d, out);
    if (!ok || rc) {
        hipFree(rows);
        return ok ? rc : PZ_FAIL(PZ_ERR_HIP, "seed expansion or transform of the Ksk rows failed");
    }
    return PZ_OK;
}

extern "C" int pz_ni_aggregate_rgsw(pz_ctx *c, const uint8_t *crs, uint32_t total,
                                    pz_rlwe_ksk *const *ksks, uint32_t l0, uint32_t count,
                                    const uint64_t *d_zero, const uint64_t *d_cts, uint64_t *d_out,
                                    void *stream)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!crs || !ksks || !d_zero || !d_cts || !d_out)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !crs || !ksks || !d_zero || !d_cts || !d_out");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (p.uitos_d == 0 || p.rgrg_da == 0 || p.rgrg_db == 0)
        return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.uitos_d == 0 || p.rgrg_da == 0 || p.rgrg_db == 0");
    // the leader's rows are the last of each part (the reference requires
    // RLWE x RGSW counts below the RGSW x RGSW ones, src/bool/evaluator.rs:1707-1716)
    if (p.rlrg_da > p.rgrg_da || p.rlrg_db > p.rgrg_db)
        return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.rlrg_da > p.rgrg_da || p.rlrg_db > p.rgrg_db");
    if (total == 0 || total > PZ_NI_MAX_PARTIES)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: total == 0 || total > PZ_NI_MAX_PARTIES");
    if ((uint64_t)l0 + count > p.lwe_n)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: (uint64_t)l0 + count > p.lwe_n");
    for (u32 o = 0; o < total; o++)
        if (!ksks[o] || !pz_internal_ctx_compatible(c, ksks[o]->owner))
            return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !ksks[o] || !pz_internal_ctx_compatible(c, ksks[o]->owner)");
    if (count == 0) return PZ_OK;
    if (stream) st = (hipStream_t)stream;
    const u32 N = p.rlwe_n, da = p.rgrg_da, db = p.rgrg_db, dmax = da > db ? da : db;
    const size_t r1 = 2 * (size_t)(da + db);
    // workspace: the others' RGSWs and their Shoup twins, a_{i,l} of the
    // RLWE'(m) rows (the last db rows of each index's stream), stream keys
    const size_t w1 = (size_t)(total - 1) * count * r1 * N, wa = (size_t)count * db * N;
    void *ws = nullptr;
    if ((rc = pz_internal_workspace(c, 8 * (2 * w1 + wa) + 48 * (size_t)count, &ws))) return rc;
    u64 *o1 = (u64 *)ws, *o1s = o1 + w1, *d_a = o1s + w1;
    std::vector<uint8_t> seeds(32 * (size_t)count);
    host::ni_index_seeds(crs, l0, count, seeds.data());
    bool ok = pz_internal_expand_uniform(st, seeds.data(), count, (u64)dmax * N,
                                         (u64)(dmax - db) * N, p.rlwe_q, d_a,
                                         (u32 *)(d_a + wa));
    if (ok) {
        NiArgs A;
        A.zero = d_zero;
        A.cts = d_cts;
        A.arows = d_a;
        for (u32 o = 0; o < PZ_NI_MAX_PARTIES; o++) {
            A.K[o] = o < total ? ksks[o]->rows : nullptr;
            A.Ks[o] = o < total ? ksks[o]->rows_s : nullptr;
        }
        A.out0 = d_out;
        A.out1 = o1;
        A.P = total;
        A.count = count;
        A.l0 = l0;
        A.per = host::ni_per_party(p.lwe_n, total);
        A.da = da;
        A.db = db;
        A.dmax = dmax;
        A.da0 = p.rlrg_da;
        A.db0 = p.rlrg_db;
        A.d = p.uitos_d;
        // consumers of one input polynomial share its digit transforms, up to
        // PZ_NI_GROUP per work group
        u32 per_t = 0;
        for (u32 r = 0; r < da + db; r++) {
            const bool lu = r < da ? r >= da - p.rlrg_da : r - da >= db - p.rlrg_db;
            per_t += (total - 1 + (lu ? 1 : 0) + PZ_NI_GROUP - 1) / PZ_NI_GROUP;
        }
        const size_t work = (size_t)count * per_t;
        const int grid = (int)(work < 65535 ? work : 65535);
        const size_t lds = 8 * (size_t)(2 * ((1 << logn) + (1 << logn) / 8));
        if (logn == 11)
            hipLaunchKernelGGL((k_ni_rows<11, PZ_NI_GROUP>), grid, 1 << 8, lds, st, A, dc);
        else
            hipLaunchKernelGGL((k_ni_rows<9, PZ_NI_GROUP>), grid, 1 << 6, lds, st, A, dc);
        if (w1) {
            launch_shoup(o1, o1s, w1, p.rlwe_q, st);
            const DevCtx cc = rgrg_devctx(p, dc);
            for (u32 s = 0; s + 1 < total; s++)
                launch_rgsw_rows(logn, d_out, 2 * (p.rlrg_da + p.rlrg_db), p.rlrg_da, p.rlrg_db,
                                 count, cc, o1 + (size_t)s * count * r1 * N,
                                 o1s + (size_t)s * count * r1 * N, (u32)r1, st);
        }
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    }
    return ok ? PZ_OK : PZ_FAIL(PZ_ERR_HIP, "aggregation launch failed");
}

// ---- automorphism keys and LWE KSK of the non-interactive server key ----

namespace pz {

// out = sum of `total` share arrays of `words` entries mod 2^k (mask)
__global__ void k_ni_sum_pow2(u64 *out, const u64 *shares, u32 total, size_t words, u64 mask)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 v = 0;
        for (u32 t = 0; t < total; t++) v += shares[(size_t)t * words + i];
        out[i] = v & mask;
    }
}

// automorphism key k: rows [a_0..a_{d-1} | b_0..b_{d-1}] of N words, the a
// rows taken in order from the expanded auto-key stream
__global__ void k_ni_auto_rows(u64 *key, const u64 *a, const u64 *b, u32 K, u32 d, u32 N)
{
    const size_t total = (size_t)K * 2 * d * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const size_t row = t / N, i = t % N;
        const size_t k = row / (2 * d), r = row % (2 * d);
        key[t] = r < d ? a[(k * d + r) * N + i] : b[(k * d + r - d) * N + i];
    }
}

// LWE-KSK rows [b_r, a_r[0..n)]
__global__ void k_ni_ksk_rows(u64 *rows, const u64 *a, const u64 *b, size_t nrows, u32 n)
{
    const size_t n1 = (size_t)n + 1, total = nrows * n1;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const size_t r = t / n1, c = t % n1;
        rows[t] = c == 0 ? b[r] : a[r * n + c - 1];
    }
}

}  // namespace pz

// The automorphism-key and LWE-KSK parts of
// aggregate_non_interactive_multi_party_server_key_shares
// (src/bool/evaluator.rs:1574-1612): the parties' b parts summed, mod Q and
// mod q_ks.
extern "C" int pz_ni_aggregate_auto_ksk(pz_ctx *c, uint32_t total,
                                        const uint64_t *d_auto_shares,
                                        const uint64_t *d_ksk_shares, uint64_t *d_auto_b,
                                        uint64_t *d_ksk_b, void *stream)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!d_auto_shares || !d_ksk_shares || !d_auto_b || !d_ksk_b)
        return PZ_FAIL(PZ_ERR_ARG, "null argument");
    if (total == 0 || total > PZ_NI_MAX_PARTIES)
        return PZ_FAIL(PZ_ERR_ARG, "total must be 1..PZ_NI_MAX_PARTIES");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (stream) st = (hipStream_t)stream;
    const size_t aw = (size_t)(p.w + 1) * p.auto_d * p.rlwe_n, kw = (size_t)p.rlwe_n * p.lwe_d;
    if (hipMemcpyAsync(d_auto_b, d_auto_shares, 8 * aw, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipMemcpyAsync failed");
    if (total > 1)
        hipLaunchKernelGGL(k_ni_sum_rows, ni_grid(aw), 256, 0, st, d_auto_b, d_auto_shares + aw,
                           total - 1, aw, p.rlwe_q);
    hipLaunchKernelGGL(k_ni_sum_pow2, ni_grid(kw), 256, 0, st, d_ksk_b, d_ksk_shares, total, kw,
                       (u64)(p.lwe_q - 1));
    return PZ_LAUNCH_STATUS();
}

// NonInteractiveServerKeyEvaluationDomain::from(&SeededNonInteractive
// MultiPartyServerKey) (src/bool/keys.rs:851-945) with the Shoup twins of
// ShoupNonInteractiveServerKeyEvaluationDomain (:1088-1155): the RGSWs
// (coefficient domain) sent to the evaluation domain; each automorphism key's
// a rows expanded from ChaCha8(puncture(puncture(crs, 1), 2)) in dlog order,
// its b rows copied, every row sent to the evaluation domain; each LWE-KSK
// row's n a entries expanded from ChaCha8(puncture(puncture(crs, 1), 3)) in
// row order after its b.  The caller's buffers are left unchanged.
extern "C" int pz_ni_expand_server_key(pz_ctx *c, const uint8_t *crs, const uint64_t *d_rgsw,
                                       const uint64_t *d_auto_b, const uint64_t *d_ksk_b,
                                       pz_key **out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!crs || !d_rgsw || !d_auto_b || !d_ksk_b || !out)
        return PZ_FAIL(PZ_ERR_ARG, "null argument");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    const u32 N = p.rlwe_n, n = p.lwe_n, d = p.auto_d, K = p.w + 1;
    const size_t rg_rows = (size_t)n * (2 * p.rlrg_da + 2 * p.rlrg_db);
    const size_t au_rows = (size_t)K * 2 * d, krows = (size_t)N * p.lwe_d;
    uint8_t sd_auto[32], sd_ksk[32];
    host::ni_seed_key_part(crs, 2, sd_auto);
    host::ni_seed_key_part(crs, 3, sd_ksk);
    u64 *rg = nullptr, *au = nullptr, *au_a = nullptr, *kr = nullptr, *ka = nullptr;
    u32 *kf = nullptr;
    bool ok = hipMalloc(&rg, 8 * rg_rows * N) == hipSuccess &&
              hipMalloc(&au, 8 * au_rows * N) == hipSuccess &&
              hipMalloc(&au_a, 8 * (size_t)K * d * N) == hipSuccess &&
              hipMalloc(&kr, 8 * krows * (n + 1)) == hipSuccess &&
              hipMalloc(&ka, 8 * krows * n) == hipSuccess && hipMalloc(&kf, 4 * 12) == hipSuccess;
    const char *what = "hipMalloc failed";
    if (ok) {
        what = "expansion of the seeded parts failed";
        ok = hipMemcpyAsync(rg, d_rgsw, 8 * rg_rows * N, hipMemcpyDeviceToDevice, st) ==
                 hipSuccess &&
             pz_internal_expand_uniform(st, sd_auto, 1, (u64)K * d * N, 0, p.rlwe_q, au_a, kf) &&
             pz_internal_expand_uniform(st, sd_ksk, 1, (u64)krows * n, 0, p.lwe_q, ka, kf);
    }
    if (ok) {
        launch_ntt(logn, 1, rg, rg_rows, dc.tb, st);
        hipLaunchKernelGGL(k_ni_auto_rows, ni_grid(au_rows * N), 256, 0, st, au, au_a, d_auto_b,
                           K, d, N);
        launch_ntt(logn, 1, au, au_rows, dc.tb, st);
        hipLaunchKernelGGL(k_ni_ksk_rows, ni_grid(krows * (n + 1)), 256, 0, st, kr, ka, d_ksk_b,
                           krows, n);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
        what = "transform or assembly of the key rows failed";
    }
    if (ok) {
        rc = pz_internal_key_adopt(c, rg, au, kr, out);  // owns rg and au on success
        if (rc == PZ_OK) rg = au = nullptr;
    } else {
        rc = PZ_FAIL(PZ_ERR_HIP, what);
    }
    hipFree(rg);
    hipFree(au);
    hipFree(au_a);
    hipFree(kr);
    hipFree(ka);
    hipFree(kf);
    return rc;
}
