// pz_input.hip -- input path of the non-interactive protocol on the device:
// RLWE key switching from a party's secret u_j to the ideal secret s, and
// sample extraction of one LWE per bit; and the RGSW x RGSW product of the
// server-key aggregation.
//
//   k_rlwe_ksw         rlwe_key_switch (src/rgsw/runtime.rs:1001-1063), one
//                      RLWE per work group: -a decomposed into d digit rows
//                      (DecomposerIter order), each row transformed and
//                      multiply-accumulated into A' and B' (rows of Ksk), one
//                      strict inverse per output, then B' + b
//   k_sample_extract   sample_extract (src/pbs.rs:423-449) at any index
#include "pz_host.hpp"
#include "pz_internal.hpp"

namespace pz {

template <int LOGN>
__global__ __launch_bounds__(1 << (LOGN - 3)) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_rlwe_ksw(const u64 *in, u64 *out, size_t B, DevCtx c, const u64 *K, const u64 *Ks, u32 d)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ u64 lds[];
    const u32 j = threadIdx.x;
    const u64 q = c.tb.q, q2 = c.tb.q2;
    for (size_t bb = blockIdx.x; bb < B; bb += gridDim.x) {
        const u64 *p = in + bb * 2 * N;
        u64 val[8];
#pragma unroll
        for (int r = 0; r < 8; r++) val[r] = decomp_prepare(c.dec_uitos, neg_mod(p[coef_x<LOGN>(j, r)], q));
        u64 acc[2][8];
#pragma unroll
        for (int r = 0; r < 8; r++) acc[0][r] = acc[1][r] = 0;
        for (u32 row = 0; row < d; row++) {
            __builtin_amdgcn_sched_barrier(0);  // no key loads hoisted across rows
            u64 D[1][8];
#pragma unroll
            for (int r = 0; r < 8; r++) D[0][r] = decomp_next(c.dec_uitos, val[r]);
            ntt_fwd<LOGN, 1, PZ_FWD_MODE, false>(D, j, lds, c.tb);
            fma_row<false>(acc[0], K + (size_t)row * N, Ks + (size_t)row * N, D[0], j, q, q2);
            fma_row<false>(acc[1], K + (size_t)(d + row) * N, Ks + (size_t)(d + row) * N, D[0], j,
                          q, q2);
        }
        ntt_inv<LOGN, 2, true, true, false>(acc, j, lds, c.tb);
        u64 *o = out + bb * 2 * N;
        u64 bv[8];
#pragma unroll
        for (int r = 0; r < 8; r++) bv[r] = p[N + coef_x<LOGN>(j, r)];
        __syncthreads();  // out may alias in: every read of this RLWE is done
#pragma unroll
        for (int r = 0; r < 8; r++) {
            o[coef_x<LOGN>(j, r)] = acc[0][r];
            o[N + coef_x<LOGN>(j, r)] = add_mod(acc[1][r], bv[r], q);
        }
    }
}

__global__ void k_sample_extract(const u64 *rlwe, const u32 *src, size_t count, u32 N, u64 q,
                                 u64 *lwe)
{
    const size_t N1 = (size_t)N + 1;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < count * N1;
         t += (size_t)gridDim.x * blockDim.x) {
        const size_t o = t / N1;
        const u32 col = (u32)(t % N1);
        const u64 *r = rlwe + (size_t)src[2 * o] * 2 * N;
        const u32 k = src[2 * o + 1];
        u64 v;
        if (k >= N) {
            v = 0;  // no such coefficient: an all-zero row, nothing read
        } else if (col == 0) {
            v = r[N + k];                             // b at the index
        } else {
            const u32 i = col - 1;                    // a_i of the LWE
            v = i <= k ? r[k - i] : neg_mod(r[N + k - i], q);
        }
        lwe[t] = v;
    }
}

}  // namespace pz

using namespace pz;

extern int pz_internal_ctx_view(pz_ctx *c, pz_params *p, DevCtx *dc, int *logn, hipStream_t *s);
extern int pz_internal_workspace(pz_ctx *c, size_t bytes, void **out);

int pz_internal_rlwe_ksk_adopt(pz_ctx *c, u64 *d_rows, u32 d, pz_rlwe_ksk **out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    const size_t words = 2 * (size_t)d * p.rlwe_n;
    pz_rlwe_ksk *k = new pz_rlwe_ksk();
    k->owner = c;
    k->d = d;
    if (hipMalloc(&k->rows_s, 8 * words) != hipSuccess) {
        delete k;
        return PZ_FAIL(PZ_ERR_HIP, "hipMalloc of the Shoup twins failed");
    }
    launch_shoup(d_rows, k->rows_s, words, p.rlwe_q, st);
    if (hipStreamSynchronize(st) != hipSuccess) {
        hipFree(k->rows_s);
        delete k;
        return PZ_FAIL(PZ_ERR_HIP, "Shoup twins of the Ksk rows failed");
    }
    k->rows = d_rows;
    *out = k;
    return PZ_OK;
}

extern "C" int pz_rlwe_ksk_upload(pz_ctx *c, const uint64_t *h_rows, pz_rlwe_ksk **out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!h_rows || !out) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !h_rows || !out");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (p.uitos_d == 0) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.uitos_d == 0");
    const size_t words = 2 * (size_t)p.uitos_d * p.rlwe_n;
    u64 *rows = nullptr;
    if (hipMalloc(&rows, 8 * words) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
    if (pz_copy_sync(rows, h_rows, 8 * words, hipMemcpyHostToDevice, st) != hipSuccess) {
        hipFree(rows);
        return PZ_FAIL(PZ_ERR_HIP, "hipMemcpy of the Ksk rows failed");
    }
    rc = pz_internal_rlwe_ksk_adopt(c, rows, p.uitos_d, out);
    if (rc) hipFree(rows);
    return rc;
}

extern "C" int pz_rlwe_ksk_download(pz_rlwe_ksk *k, uint64_t *h_rows)
{
    if (!k || !h_rows) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !k || !h_rows");
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    int rc = pz_internal_ctx_view(k->owner, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (pz_copy_sync(h_rows, k->rows, 16 * (size_t)k->d * p.rlwe_n, hipMemcpyDeviceToHost, st) !=
        hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "copy failed");
    return PZ_OK;
}

extern "C" int pz_rlwe_ksk_free(pz_rlwe_ksk *k)
{
    if (!k) return PZ_OK;
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (pz_internal_ctx_view(k->owner, &p, &dc, &logn, &st) == PZ_OK) hipStreamSynchronize(st);
    hipFree(k->rows);
    hipFree(k->rows_s);
    delete k;
    return PZ_OK;
}

extern "C" int pz_rlwe_key_switch(pz_ctx *c, pz_rlwe_ksk *k, const uint64_t *d_in, uint64_t *d_out,
                                  size_t B, void *stream)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!k || !d_in || !d_out)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !k || !d_in || !d_out");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (p.uitos_d == 0) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.uitos_d == 0");
    if (!pz_internal_ctx_compatible(c, k->owner))
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !pz_internal_ctx_compatible(c, k->owner)");
    if (B == 0) return PZ_OK;
    if (stream) st = (hipStream_t)stream;
    const int grid = (int)(B < 65535 ? B : 65535);
    const size_t lds = 8 * (size_t)(2 * ((1 << logn) + (1 << logn) / 8));
    switch (logn) {
    case 9:
        hipLaunchKernelGGL(k_rlwe_ksw<9>, grid, 1 << 6, lds, st, d_in, d_out, B, dc, k->rows,
                           k->rows_s, k->d);
        break;
    case 11:
        hipLaunchKernelGGL(k_rlwe_ksw<11>, grid, 1 << 8, lds, st, d_in, d_out, B, dc, k->rows,
                           k->rows_s, k->d);
        break;
    default:
        return PZ_FAIL(PZ_ERR_UNSUPPORTED, "ring size must be 512 or 2048");
    }
    return PZ_LAUNCH_STATUS();
}

extern "C" int pz_sample_extract(pz_ctx *c, const uint64_t *d_rlwe, const uint32_t *d_src,
                                 size_t count, uint64_t *d_lwe, void *stream)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!d_rlwe || !d_src || !d_lwe)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !d_rlwe || !d_src || !d_lwe");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (count == 0) return PZ_OK;
    if (stream) st = (hipStream_t)stream;
    const size_t total = count * ((size_t)p.rlwe_n + 1);
    const int grid = (int)((total + 255) / 256 < 65535 * 4 ? (total + 255) / 256 : 65535 * 4);
    hipLaunchKernelGGL(k_sample_extract, grid, 256, 0, st, d_rlwe, d_src, count, p.rlwe_n,
                       p.rlwe_q, d_lwe);
    return PZ_LAUNCH_STATUS();
}

// The context with the RGSW x RGSW decomposers in place of the RLWE x RGSW
// ones (rgsw_x_rgsw_decomposer, src/bool/parameters.rs)
DevCtx pz::rgrg_devctx(const pz_params &p, const DevCtx &dc)
{
    DevCtx cc = dc;
    const u32 lq = host::ref_log2(p.rlwe_q);
    cc.dec_a.q = cc.dec_b.q = p.rlwe_q;
    cc.dec_a.logb = cc.dec_b.logb = p.rgrg_logb;
    cc.dec_a.d = p.rgrg_da;
    cc.dec_b.d = p.rgrg_db;
    cc.dec_a.ignore_bits = lq - p.rgrg_logb * p.rgrg_da;
    cc.dec_b.ignore_bits = lq - p.rgrg_logb * p.rgrg_db;
    decomp_centring(cc.dec_a);
    decomp_centring(cc.dec_b);
    cc.da = p.rgrg_da;
    cc.db = p.rgrg_db;
    return cc;
}

// RGSW x RGSW product of the non-interactive server-key aggregation
// (rgsw_by_rgsw_inplace, src/rgsw/runtime.rs:883-985, applied per LWE index
// and party at src/bool/evaluator.rs:1505-1535).
extern "C" int pz_rgsw_by_rgsw(pz_ctx *c, uint64_t *d_rgsw0, const uint64_t *d_rgsw1, size_t B,
                               void *stream)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!d_rgsw0 || !d_rgsw1) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !d_rgsw0 || !d_rgsw1");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (p.rgrg_da == 0 || p.rgrg_db == 0)
        return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.rgrg_da == 0 || p.rgrg_db == 0");
    if (B == 0) return PZ_OK;
    if (stream) st = (hipStream_t)stream;
    const u32 N = p.rlwe_n, r0 = 2 * (p.rlrg_da + p.rlrg_db), r1 = 2 * (p.rgrg_da + p.rgrg_db);
    const size_t words = B * (size_t)r1 * N;
    // RGSW1 rows times n^-1 (the step's inverse transforms leave it out) and
    // their Shoup twins, in scratch
    void *ws = nullptr;
    if ((rc = pz_internal_workspace(c, 16 * words, &ws))) return rc;
    u64 *k1 = (u64 *)ws, *k1s = k1 + words;
    hipMemcpyAsync(k1, d_rgsw1, 8 * words, hipMemcpyDeviceToDevice, st);
    launch_scale(k1, words, dc.tb.n_inv, dc.tb.n_inv_s, p.rlwe_q, st);
    launch_shoup(k1, k1s, words, p.rlwe_q, st);
    const DevCtx cc = rgrg_devctx(p, dc);
    launch_rgsw_rows(logn, d_rgsw0, r0, p.rlrg_da, p.rlrg_db, B, cc, k1, k1s, r1, st);
    hipError_t e = hipGetLastError();
    if (hipStreamSynchronize(st) != hipSuccess) e = hipErrorUnknown;  // workspace free again
    return e == hipSuccess ? PZ_OK : PZ_FAIL(PZ_ERR_HIP, hipGetErrorString(e));
}

