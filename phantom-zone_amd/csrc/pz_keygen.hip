// pz_keygen.hip -- seeded client-side key generation and LWE encryption.
//
// Secrets and all random polynomials are drawn on the host from ChaCha8
// streams (the stream order is the project's own, fixed below); the products
// a*s, the message terms and the transforms to the evaluation domain run on
// the device with the same exact arithmetic as the reference's keygen:
//   RGSW(X^{emb*z_i})   secret_key_encrypt_rgsw (src/rgsw/keygen.rs:68-160),
//                       both RLWE'(m) halves explicit, then forward NTT of
//                       every row (src/bool/keys.rs:881-885);
//   auto keys           seeded_auto_key_gen / seeded_rlwe_ksk_gen
//                       (src/rgsw/keygen.rs:306-430) for [-g, g, .., g^w];
//   LWE KSK             seeded_lwe_ksk_keygen (src/lwe.rs:46-101).
// Stream order (seed_key): for each RGSW i, for each of its d_a + d_b row
// pairs: N uniform a, N noise e; then the auto keys likewise; then per KSK
// row: n uniform a (mod q_ks), one noise e.
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/pz.h"
#include "pz_host.hpp"
#include "pz_internal.hpp"

namespace pz {

// One RLWE row pair of key material.
struct RowPair {
    const u64 *a;    // uniform a (coefficient domain)
    const u64 *e;    // noise, already mod q
    const u64 *as;   // a * s
    const u64 *msg;  // full message polynomial (auto keys) or null
    u64 *dst_a, *dst_b;
    u64 beta;
    u32 mono_pos;    // monomial message position (RGSW rows)
    u64 mono_val;    // 1 or q - 1; 0 = no monomial
    u32 msg_in_a;    // 1: message added to the a row (RLWE'(-sm)), 0: to b
};

__global__ void k_barrett_rows(u64 *rows, const u64 *s_eval, size_t nrows, u32 N, Barrett br)
{
    size_t total = nrows * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x)
        rows[i] = barrett_mul(br, rows[i], s_eval[i % N]);
}

// b = (e + a*s) + beta*msg_b ; a' = a + beta*msg_a  (add_mod_fast /
// mul_mod_fast, src/backend/modulus_u64.rs:46-55, :106-130)
__global__ void k_kg_combine(const RowPair *pairs, size_t npairs, u32 N, Barrett br)
{
    const u64 q = br.q;
    size_t total = npairs * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const RowPair &p = pairs[t / N];
        u32 i = (u32)(t % N);
        u64 m = 0;
        if (p.msg) m = p.msg[i];
        else if (p.mono_val && i == p.mono_pos) m = p.mono_val;
        u64 bm = barrett_mul(br, m, p.beta);
        u64 a = p.a[i];
        u64 b = add_mod(p.e[i], p.as[i], q);
        if (p.msg_in_a) a = add_mod(a, bm, q);
        else b = add_mod(b, bm, q);
        p.dst_a[i] = a;
        p.dst_b[i] = b;
    }
}

// KSK row r = i*d + j: b = <a, z> + beta_j * (-s_i) + e (mod q_ks)
__global__ void k_ksk_rows(u64 *ksk, const int32_t *s, const int32_t *z, const int64_t *err,
                           u32 N, u32 d, u32 n, u32 logb, u32 ib, u64 mask)
{
    size_t rows = (size_t)N * d;
    size_t n1 = (size_t)n + 1;
    for (size_t r = blockIdx.x * (size_t)blockDim.x + threadIdx.x; r < rows;
         r += (size_t)gridDim.x * blockDim.x) {
        u64 *row = ksk + r * n1;
        u64 az = 0;
        for (u32 k = 0; k < n; k++) az += row[1 + k] * (u64)(int64_t)z[k];
        u32 i = (u32)(r / d), j = (u32)(r % d);
        u64 beta = 1ull << (logb * j + ib);
        u64 neg_si = 0 - (u64)(int64_t)s[i];
        u64 b = (az + beta * neg_si) & mask;
        b = (b + (u64)err[r]) & mask;
        row[0] = b;
    }
}

// NI share rows: out = (prod + e) + beta * msg, msg the party's u (u -> s
// rows), a monomial X^{emb*z_j[l]} (RGSW rows) or nothing (zero encryptions)
struct NiRow {
    const u64 *prod;  // a*s_j or a*u_j
    const u64 *e;
    const u64 *msg;   // full message polynomial or null
    u64 *out;
    u64 beta;
    u32 mono_pos;
    u64 mono_val;     // 1 or q - 1; 0 = no monomial
};

__global__ void k_ni_combine(const NiRow *rows, size_t nrows, u32 N, Barrett br)
{
    const u64 q = br.q;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nrows * N;
         t += (size_t)gridDim.x * blockDim.x) {
        const NiRow &r = rows[t / N];
        const u32 i = (u32)(t % N);
        u64 m = 0;
        if (r.msg) m = r.msg[i];
        else if (r.mono_val && i == r.mono_pos) m = r.mono_val;
        r.out[i] = add_mod(add_mod(r.prod[i], r.e[i], q), barrett_mul(br, m, r.beta), q);
    }
}

// NI LWE-KSK share row r = i*d + j (seeded_lwe_ksk_keygen, src/lwe.rs:46-101):
// b[r] = <a_r, z> + beta_j * (-s_i) + e_r mod q_ks, a_r the n entries of row r
// (rows of the CRS stream).  One wave per row: lanes sum strided entries,
// then a butterfly reduction (all arithmetic wraps mod 2^64, masked at the
// end, so the order of the sum does not matter).
__global__ __launch_bounds__(64) void k_ksk_share_b(const u64 *a, const int32_t *s,
                                                    const int32_t *z, const int64_t *err, u32 N,
                                                    u32 d, u32 n, u32 logb, u32 ib, u64 mask,
                                                    u64 *b)
{
    const size_t rows = (size_t)N * d;
    for (size_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const u64 *ar = a + r * n;
        u64 acc = 0;
        for (u32 k = threadIdx.x; k < n; k += 64) acc += ar[k] * (u64)(int64_t)z[k];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (threadIdx.x == 0) {
            const u32 i = (u32)(r / d), j = (u32)(r % d);
            const u64 beta = 1ull << (logb * j + ib);
            const u64 neg_si = 0 - (u64)(int64_t)s[i];
            b[r] = (acc + beta * neg_si + (u64)err[r]) & mask;
        }
    }
}

static int grid_of(size_t n) { return (int)((n + 255) / 256 < 65535 * 4 ? (n + 255) / 256 : 65535 * 4); }

}  // namespace pz

using namespace pz;

// defined in pz_capi.hip
extern int pz_internal_ctx_view(pz_ctx *c, pz_params *p, DevCtx *dc, int *logn, hipStream_t *s);
extern int pz_internal_key_adopt(pz_ctx *c, u64 *rgsw, u64 *auto_k, const u64 *ksk_rows,
                                 pz_key **out);

extern "C" int pz_keygen_secrets(pz_ctx *c, const uint8_t *seed, uint32_t parties,
                                 int32_t *h_s, int32_t *h_z)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!seed || !h_s || !h_z || parties == 0)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !seed || !h_s || !h_z || parties == 0");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    host::keygen_secrets(p, seed, parties, h_s, h_z);
    return PZ_OK;
}

#define KG_CHK(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) {                  \
            PZ_FAIL(PZ_ERR_HIP, #x);              \
            ok = false;                           \
            goto done;                            \
        }                                         \
    } while (0)
#define KG_FAIL(what)                             \
    do {                                          \
        PZ_FAIL(PZ_ERR_HIP, what);                \
        ok = false;                               \
        goto done;                                \
    } while (0)

extern "C" int pz_keygen_server(pz_ctx *c, const uint8_t *seed, const int32_t *h_s,
                                const int32_t *h_z, pz_key **out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!seed || !h_s || !h_z || !out)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !seed || !h_s || !h_z || !out");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    const u64 q = p.rlwe_q;
    const u32 N = p.rlwe_n, n = p.lwe_n, da = p.rlrg_da, db = p.rlrg_db, dau = p.auto_d;
    const u32 emb = (2 * N) / p.br_q;
    const size_t rgsw_rows = 2 * (size_t)da + 2 * db, auto_rows = 2 * (size_t)dau;
    const size_t n_rg_pairs = (size_t)n * (da + db), n_au_pairs = (size_t)(p.w + 1) * dau;
    const size_t npairs = n_rg_pairs + n_au_pairs;
    const size_t ksk_rows = (size_t)N * p.lwe_d, n1 = (size_t)n + 1;
    host::ChaCha8 r(seed);

    // ---- host sampling, in stream order ----
    std::vector<u64> ha(npairs * N), he(npairs * N);
    for (size_t pr = 0; pr < npairs; pr++) {
        for (u32 i = 0; i < N; i++) ha[pr * N + i] = r.uniform(q);
        for (u32 i = 0; i < N; i++) he[pr * N + i] = host::map_i64(q, host::gauss_round(r));
    }
    std::vector<u64> hksk(ksk_rows * n1);
    std::vector<int64_t> herr(ksk_rows);
    for (size_t row = 0; row < ksk_rows; row++) {
        for (u32 k = 0; k < n; k++) hksk[row * n1 + 1 + k] = r.uniform(p.lwe_q);
        herr[row] = (int64_t)host::map_i64(p.lwe_q, host::gauss_round(r));
    }
    // auto-key messages -s(X^k) (seeded_auto_key_gen, src/rgsw/keygen.rs:403-415)
    std::vector<u64> hmsg((size_t)(p.w + 1) * N);
    for (u32 k = 0; k <= p.w; k++) {
        uint64_t kk = dc.auto_el[k];
        for (u32 i = 0; i < N; i++) {
            uint64_t to = ((uint64_t)i * kk) % (2ull * N);
            bool pos = to < N;
            if (!pos) to -= N;
            u64 el = host::map_i64(q, h_s[i]);
            hmsg[(size_t)k * N + to] = pos ? q - el : el;
        }
    }
    std::vector<u64> hs_mod(N);
    for (u32 i = 0; i < N; i++) hs_mod[i] = host::map_i64(q, h_s[i]);

    bool ok = true;
    u64 *d_a = nullptr, *d_e = nullptr, *d_as = nullptr, *d_s = nullptr, *d_msg = nullptr;
    u64 *d_rgsw = nullptr, *d_auto = nullptr, *d_ksk = nullptr;
    int32_t *d_si = nullptr, *d_zi = nullptr;
    int64_t *d_err = nullptr;
    RowPair *d_pairs = nullptr;
    std::vector<RowPair> pairs(npairs);
    const u32 lq_log = host::ref_log2(p.lwe_q);
    const u32 ib_lwe = lq_log - p.lwe_logb * p.lwe_d;
    KG_CHK(hipMalloc(&d_a, 8 * npairs * N));
    KG_CHK(hipMalloc(&d_e, 8 * npairs * N));
    KG_CHK(hipMalloc(&d_as, 8 * npairs * N));
    KG_CHK(hipMalloc(&d_s, 8 * (size_t)N));
    KG_CHK(hipMalloc(&d_msg, 8 * hmsg.size()));
    KG_CHK(hipMalloc(&d_rgsw, 8 * (size_t)n * rgsw_rows * N));
    KG_CHK(hipMalloc(&d_auto, 8 * (size_t)(p.w + 1) * auto_rows * N));
    KG_CHK(hipMalloc(&d_ksk, 8 * ksk_rows * n1));
    KG_CHK(hipMalloc(&d_si, 4 * (size_t)N));
    KG_CHK(hipMalloc(&d_zi, 4 * (size_t)n));
    KG_CHK(hipMalloc(&d_err, 8 * ksk_rows));
    KG_CHK(hipMalloc(&d_pairs, sizeof(RowPair) * npairs));
    KG_CHK(hipMemcpyAsync(d_a, ha.data(), 8 * ha.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_e, he.data(), 8 * he.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_as, ha.data(), 8 * ha.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_s, hs_mod.data(), 8 * (size_t)N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_msg, hmsg.data(), 8 * hmsg.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_ksk, hksk.data(), 8 * hksk.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_si, h_s, 4 * (size_t)N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_zi, h_z, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_err, herr.data(), 8 * ksk_rows, hipMemcpyHostToDevice, st));

    // s_eval, then a*s = backward(forward(a) . s_eval) for every row pair
    launch_ntt(logn, 1, d_s, 1, dc.tb, st);
    launch_ntt(logn, 1, d_as, npairs, dc.tb, st);
    hipLaunchKernelGGL(k_barrett_rows, grid_of(npairs * N), 256, 0, st, d_as, d_s, npairs, N,
                       dc.br);
    launch_ntt(logn, 3, d_as, npairs, dc.tb, st);

    // row-pair descriptors
    for (u32 i = 0; i < n; i++) {
        int64_t si = (int64_t)emb * h_z[i];
        u32 pos = si < 0 ? (u32)(N - (u32)(-si)) : (u32)si;
        u64 val = si < 0 ? q - 1 : 1;
        u64 *ct = d_rgsw + (size_t)i * rgsw_rows * N;
        for (u32 j = 0; j < da + db; j++) {
            size_t pr = (size_t)i * (da + db) + j;
            RowPair &rp = pairs[pr];
            rp.a = d_a + pr * N;
            rp.e = d_e + pr * N;
            rp.as = d_as + pr * N;
            rp.msg = nullptr;
            rp.mono_pos = pos;
            rp.mono_val = val;
            if (j < da) {
                rp.beta = 1ull << (p.rlrg_logb * j + dc.dec_a.ignore_bits);
                rp.dst_a = ct + (size_t)j * N;
                rp.dst_b = ct + (size_t)(da + j) * N;
                rp.msg_in_a = 1;
            } else {
                u32 jj = j - da;
                rp.beta = 1ull << (p.rlrg_logb * jj + dc.dec_b.ignore_bits);
                rp.dst_a = ct + (size_t)(2 * da + jj) * N;
                rp.dst_b = ct + (size_t)(2 * da + db + jj) * N;
                rp.msg_in_a = 0;
            }
        }
    }
    for (u32 k = 0; k <= p.w; k++) {
        u64 *ct = d_auto + (size_t)k * auto_rows * N;
        for (u32 j = 0; j < dau; j++) {
            size_t pr = n_rg_pairs + (size_t)k * dau + j;
            RowPair &rp = pairs[pr];
            rp.a = d_a + pr * N;
            rp.e = d_e + pr * N;
            rp.as = d_as + pr * N;
            rp.msg = d_msg + (size_t)k * N;
            rp.mono_pos = 0;
            rp.mono_val = 0;
            rp.beta = 1ull << (p.auto_logb * j + dc.dec_auto.ignore_bits);
            rp.dst_a = ct + (size_t)j * N;
            rp.dst_b = ct + (size_t)(dau + j) * N;
            rp.msg_in_a = 0;
        }
    }
    KG_CHK(hipMemcpyAsync(d_pairs, pairs.data(), sizeof(RowPair) * npairs, hipMemcpyHostToDevice,
                          st));
    hipLaunchKernelGGL(k_kg_combine, grid_of(npairs * N), 256, 0, st, d_pairs, npairs, N, dc.br);
    // evaluation domain (Ntt::forward of every row)
    launch_ntt(logn, 1, d_rgsw, (size_t)n * rgsw_rows, dc.tb, st);
    launch_ntt(logn, 1, d_auto, (size_t)(p.w + 1) * auto_rows, dc.tb, st);
    hipLaunchKernelGGL(k_ksk_rows, grid_of(ksk_rows), 256, 0, st, d_ksk, d_si, d_zi, d_err, N,
                       p.lwe_d, n, p.lwe_logb, ib_lwe, (u64)(p.lwe_q - 1));
    KG_CHK(hipGetLastError());
    KG_CHK(hipStreamSynchronize(st));
    rc = pz_internal_key_adopt(c, d_rgsw, d_auto, d_ksk, out);
    if (rc == PZ_OK) {
        d_rgsw = nullptr;  // owned by the key now
        d_auto = nullptr;
    }
done:
    hipFree(d_a);
    hipFree(d_e);
    hipFree(d_as);
    hipFree(d_s);
    hipFree(d_msg);
    hipFree(d_rgsw);
    hipFree(d_auto);
    hipFree(d_ksk);
    hipFree(d_si);
    hipFree(d_zi);
    hipFree(d_err);
    hipFree(d_pairs);
    if (!ok) return PZ_ERR_HIP;  // message set by KG_CHK / KG_FAIL
    return rc;
}

// ---- client-side LWE encryption / decryption (encrypt_lwe / decrypt_lwe,
// src/lwe.rs:104-162; decode, src/bool/evaluator.rs:245-258).  Ciphertext i
// of a batch uses the seed (seed + i) read as a 256-bit little-endian
// integer: a = dim uniform draws, then one noise draw.
static void seed_plus(const uint8_t *seed, uint64_t add, uint8_t out[32])
{
    unsigned carry = 0;
    for (int b = 0; b < 32; b++) {
        unsigned v = seed[b] + (unsigned)((b < 8 ? (add >> (8 * b)) : 0) & 0xff) + carry;
        out[b] = (uint8_t)v;
        carry = v >> 8;
    }
}

extern "C" int pz_encrypt_bools(pz_ctx *c, const int32_t *h_s, uint32_t dim, const uint8_t *seed,
                                const uint8_t *h_bits, size_t count, uint64_t *h_out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!h_s || !seed || !h_bits || !h_out)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !h_s || !seed || !h_bits || !h_out");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    const u64 q = p.rlwe_q, tru = q >> 3, fal = (q - 1) - tru + 1;
    for (size_t t = 0; t < count; t++) {
        uint8_t sd[32];
        seed_plus(seed, t, sd);
        host::ChaCha8 r(sd);
        uint64_t *o = h_out + t * (dim + 1);
        host::u128 sa = 0;
        for (uint32_t i = 0; i < dim; i++) o[1 + i] = r.uniform(q);
        uint64_t acc = 0;
        for (uint32_t i = 0; i < dim; i++) {
            acc = (uint64_t)(((host::u128)o[1 + i] * host::map_i64(q, h_s[i]) + acc) % q);
        }
        (void)sa;
        uint64_t e = host::map_i64(q, host::gauss_round(r));
        uint64_t m = h_bits[t] ? tru : fal;
        uint64_t b = acc + e;
        b = b >= q ? b - q : b;
        b += m;
        o[0] = b >= q ? b - q : b;
    }
    return PZ_OK;
}

extern "C" int pz_decrypt_bools(pz_ctx *c, const int32_t *h_s, uint32_t dim, const uint64_t *h_cts,
                                size_t count, uint8_t *h_bits)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!h_s || !h_cts || !h_bits)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !h_s || !h_cts || !h_bits");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    const u64 q = p.rlwe_q, qby8 = q >> 3;
    int bad = 0;
    for (size_t t = 0; t < count; t++) {
        const uint64_t *ct = h_cts + t * (dim + 1);
        uint64_t acc = 0;
        for (uint32_t i = 0; i < dim; i++)
            acc = (uint64_t)(((host::u128)ct[1 + i] * host::map_i64(q, h_s[i]) + acc) % q);
        uint64_t m = ct[0] >= acc ? ct[0] - acc : q + ct[0] - acc;
        double x = ((double)(m + qby8) * 4.0) / (double)q;
        uint64_t v = ((uint64_t)__builtin_round(x)) % 4;
        if (v > 1) bad = 1;
        h_bits[t] = (uint8_t)(v == 1);
    }
    return bad ? PZ_FAIL(PZ_ERR_STATE, "a ciphertext decodes to neither 0 nor 1") : PZ_OK;
}

// ---- input path of the non-interactive protocol (the library's own seeded
// procedures; structure of ClientKey::sk_u_rlwe, ui_to_s_ksk and the batched
// bool encryption, src/bool/keys.rs:75-154, src/bool/ni_mp_api.rs:294-352,
// seeded_secret_key_encrypt_rlwe src/rgsw/keygen.rs:439-481).
extern int pz_internal_rlwe_ksk_adopt(pz_ctx *c, u64 *d_rows, u32 d, pz_rlwe_ksk **out);

extern "C" int pz_keygen_user_secret(pz_ctx *c, const uint8_t *seed, int32_t *h_u)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!seed || !h_u) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !seed || !h_u");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    host::ChaCha8 r(seed);
    host::ternary_hw(r, h_u, p.rlwe_n, p.rlwe_n >> 1);
    return PZ_OK;
}

// Ksk(u -> s): per row j, N uniform a then N noise draws (stream order);
// B_j = (e + a*s) + 2^(logb*j + ib)*u, rows [A_0..A_{d-1}, B_0..B_{d-1}],
// then Ntt::forward of every row
extern "C" int pz_keygen_rlwe_ksk(pz_ctx *c, const uint8_t *seed, const int32_t *h_u,
                                  const int32_t *h_s, pz_rlwe_ksk **out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!seed || !h_u || !h_s || !out)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !seed || !h_u || !h_s || !out");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (p.uitos_d == 0) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.uitos_d == 0");
    const u64 q = p.rlwe_q;
    const u32 N = p.rlwe_n, d = p.uitos_d;
    host::ChaCha8 r(seed);
    std::vector<u64> ha((size_t)d * N), he((size_t)d * N), hs(N), hu(N);
    for (u32 j = 0; j < d; j++) {
        for (u32 i = 0; i < N; i++) ha[(size_t)j * N + i] = r.uniform(q);
        for (u32 i = 0; i < N; i++) he[(size_t)j * N + i] = host::map_i64(q, host::gauss_round(r));
    }
    for (u32 i = 0; i < N; i++) {
        hs[i] = host::map_i64(q, h_s[i]);
        hu[i] = host::map_i64(q, h_u[i]);
    }
    bool ok = true;
    u64 *d_a = nullptr, *d_e = nullptr, *d_as = nullptr, *d_s = nullptr, *d_u = nullptr;
    u64 *d_rows = nullptr;
    RowPair *d_pairs = nullptr;
    std::vector<RowPair> pairs(d);
    KG_CHK(hipMalloc(&d_a, 8 * (size_t)d * N));
    KG_CHK(hipMalloc(&d_e, 8 * (size_t)d * N));
    KG_CHK(hipMalloc(&d_as, 8 * (size_t)d * N));
    KG_CHK(hipMalloc(&d_s, 8 * (size_t)N));
    KG_CHK(hipMalloc(&d_u, 8 * (size_t)N));
    KG_CHK(hipMalloc(&d_rows, 16 * (size_t)d * N));
    KG_CHK(hipMalloc(&d_pairs, sizeof(RowPair) * d));
    KG_CHK(hipMemcpyAsync(d_a, ha.data(), 8 * ha.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_as, ha.data(), 8 * ha.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_e, he.data(), 8 * he.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_s, hs.data(), 8 * (size_t)N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_u, hu.data(), 8 * (size_t)N, hipMemcpyHostToDevice, st));
    launch_ntt(logn, 1, d_s, 1, dc.tb, st);
    launch_ntt(logn, 1, d_as, d, dc.tb, st);
    hipLaunchKernelGGL(k_barrett_rows, grid_of((size_t)d * N), 256, 0, st, d_as, d_s, (size_t)d,
                       N, dc.br);
    launch_ntt(logn, 3, d_as, d, dc.tb, st);
    for (u32 j = 0; j < d; j++) {
        RowPair &rp = pairs[j];
        rp.a = d_a + (size_t)j * N;
        rp.e = d_e + (size_t)j * N;
        rp.as = d_as + (size_t)j * N;
        rp.msg = d_u;
        rp.mono_pos = 0;
        rp.mono_val = 0;
        rp.beta = 1ull << (p.uitos_logb * j + dc.dec_uitos.ignore_bits);
        rp.dst_a = d_rows + (size_t)j * N;
        rp.dst_b = d_rows + (size_t)(d + j) * N;
        rp.msg_in_a = 0;
    }
    KG_CHK(hipMemcpyAsync(d_pairs, pairs.data(), sizeof(RowPair) * d, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_kg_combine, grid_of((size_t)d * N), 256, 0, st, d_pairs, (size_t)d, N,
                       dc.br);
    launch_ntt(logn, 1, d_rows, 2 * (size_t)d, dc.tb, st);
    KG_CHK(hipGetLastError());
    KG_CHK(hipStreamSynchronize(st));
    rc = pz_internal_rlwe_ksk_adopt(c, d_rows, d, out);
    if (rc == PZ_OK) d_rows = nullptr;  // owned by the ksk now
done:
    hipFree(d_a);
    hipFree(d_e);
    hipFree(d_as);
    hipFree(d_s);
    hipFree(d_u);
    hipFree(d_rows);
    hipFree(d_pairs);
    if (!ok) return PZ_ERR_HIP;  // message set by KG_CHK / KG_FAIL
    return rc;
}

// Bool batch under u: sub-seeds seed_a, seed_e are the first and second 32
// bytes of ChaCha8(seed); per ciphertext N uniform a from seed_a's stream and
// N noise draws from seed_e's; b = (e + m) + a*u with a*u through
// forward / Barrett pointwise / backward
extern "C" int pz_encrypt_rlwe_bools(pz_ctx *c, const int32_t *h_u, const uint8_t *seed,
                                     const uint8_t *h_bits, size_t count, uint64_t *h_out)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!h_u || !seed || !h_bits || !h_out)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !h_u || !seed || !h_bits || !h_out");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (count == 0) return PZ_OK;
    const u64 q = p.rlwe_q, tru = q >> 3, fal = (q - 1) - tru + 1;
    const u32 N = p.rlwe_n;
    const size_t nr = (count + N - 1) / N;
    uint8_t sa[32], se[32];
    {
        host::ChaCha8 master(seed);
        master.fill_bytes(sa, 32);
        master.fill_bytes(se, 32);
    }
    host::ChaCha8 ra(sa), re(se);
    std::vector<u64> ha(nr * N), hm(nr * N), hu(N);
    for (size_t k = 0; k < nr; k++) {
        for (u32 i = 0; i < N; i++) ha[k * N + i] = ra.uniform(q);
        for (u32 i = 0; i < N; i++) {
            size_t t = k * N + i;
            u64 e = host::map_i64(q, host::gauss_round(re));
            u64 m = t < count ? (h_bits[t] ? tru : fal) : 0;
            u64 v = e + m;
            hm[t] = v >= q ? v - q : v;  // e + m (add_mod_fast)
        }
    }
    for (u32 i = 0; i < N; i++) hu[i] = host::map_i64(q, h_u[i]);
    bool ok = true;
    u64 *d_as = nullptr, *d_u = nullptr;
    KG_CHK(hipMalloc(&d_as, 8 * nr * N));
    KG_CHK(hipMalloc(&d_u, 8 * (size_t)N));
    KG_CHK(hipMemcpyAsync(d_as, ha.data(), 8 * ha.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_u, hu.data(), 8 * (size_t)N, hipMemcpyHostToDevice, st));
    launch_ntt(logn, 1, d_u, 1, dc.tb, st);
    launch_ntt(logn, 1, d_as, nr, dc.tb, st);
    hipLaunchKernelGGL(k_barrett_rows, grid_of(nr * N), 256, 0, st, d_as, d_u, nr, N, dc.br);
    launch_ntt(logn, 3, d_as, nr, dc.tb, st);
    {
        std::vector<u64> has(nr * N);
        KG_CHK(hipMemcpyAsync(has.data(), d_as, 8 * has.size(), hipMemcpyDeviceToHost, st));
        KG_CHK(hipStreamSynchronize(st));
        for (size_t k = 0; k < nr; k++) {
            uint64_t *o = h_out + k * 2 * N;
            for (u32 i = 0; i < N; i++) {
                o[i] = ha[k * N + i];
                u64 v = hm[k * N + i] + has[k * N + i];
                o[N + i] = v >= q ? v - q : v;
            }
        }
    }
done:
    hipFree(d_as);
    hipFree(d_u);
    return ok ? PZ_OK : PZ_ERR_HIP;
}

// ---- non-interactive server-key shares (client side).  Uniform rows and
// noise are drawn on the host (one ChaCha8 stream per seed, several indices
// at once); products with the party's secrets run on the device.

// rows[r] (coefficient domain, r < nrows) times x (a secret) into prod:
// forward, Barrett pointwise, backward (the reference's keygen products)
static bool ni_products(int logn, const DevCtx &dc, hipStream_t st, const u64 *d_rows,
                        size_t nrows, const int32_t *h_x, u32 N, u64 q, u64 *d_prod)
{
    std::vector<u64> hx(N);
    for (u32 i = 0; i < N; i++) hx[i] = host::map_i64(q, h_x[i]);
    u64 *d_x = nullptr;
    if (hipMalloc(&d_x, 8 * (size_t)N) != hipSuccess) return false;
    bool ok = hipMemcpyAsync(d_x, hx.data(), 8 * (size_t)N, hipMemcpyHostToDevice, st) ==
                  hipSuccess &&
              hipMemcpyAsync(d_prod, d_rows, 8 * nrows * N, hipMemcpyDeviceToDevice, st) ==
                  hipSuccess;
    if (ok) {
        launch_ntt(logn, 1, d_x, 1, dc.tb, st);
        launch_ntt(logn, 1, d_prod, nrows, dc.tb, st);
        hipLaunchKernelGGL(k_barrett_rows, grid_of(nrows * N), 256, 0, st, d_prod, d_x, nrows, N,
                           dc.br);
        launch_ntt(logn, 3, d_prod, nrows, dc.tb, st);
        ok = hipStreamSynchronize(st) == hipSuccess;
    }
    hipFree(d_x);
    return ok;
}

static bool ni_combine(hipStream_t st, const DevCtx &dc, const std::vector<NiRow> &rows, u32 N)
{
    NiRow *d_rows = nullptr;
    if (rows.empty()) return true;
    if (hipMalloc(&d_rows, sizeof(NiRow) * rows.size()) != hipSuccess) return false;
    bool ok = hipMemcpyAsync(d_rows, rows.data(), sizeof(NiRow) * rows.size(),
                             hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_ni_combine, grid_of(rows.size() * N), 256, 0, st, d_rows,
                           rows.size(), N, dc.br);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    }
    hipFree(d_rows);
    return ok;
}

static int ni_check(const pz_params &p, uint32_t self, uint32_t total)
{
    if (p.uitos_d == 0 || p.rgrg_da == 0 || p.rgrg_db == 0)
        return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: p.uitos_d == 0 || p.rgrg_da == 0 || p.rgrg_db == 0");
    if (total == 0 || total > PZ_NI_MAX_PARTIES || self >= total)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: total == 0 || total > PZ_NI_MAX_PARTIES || self >= total");
    return PZ_OK;
}

extern "C" int pz_ni_share_uitos(pz_ctx *c, const uint8_t *crs, const uint8_t *noise_seed,
                                 uint32_t self, uint32_t total, const int32_t *h_s,
                                 const int32_t *h_u, uint64_t *h_ksk_b, uint64_t *h_zero_others)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!crs || !noise_seed || !h_s || !h_u || !h_ksk_b || (total > 1 && !h_zero_others))
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !crs || !noise_seed || !h_s || !h_u || !h_ksk_b || (total > 1 && !h_zero_others");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if ((rc = ni_check(p, self, total))) return rc;
    const u64 q = p.rlwe_q;
    const u32 N = p.rlwe_n, d = p.uitos_d;
    const size_t R = (size_t)total * d;  // own rows, then the others' in ascending order
    std::vector<u64> ha(R * N), he(R * N), hu(N);
    {
        uint8_t ns[32], sd[32];
        host::puncture(noise_seed, 1, ns);
        host::ChaCha8 noise(ns);
        size_t row = 0;
        for (u32 pass = 0; pass < total; pass++) {
            const u32 o = pass == 0 ? self : (pass - 1 < self ? pass - 1 : pass);
            host::ni_seed_user(crs, o, sd);
            host::ChaCha8 pr(sd);
            for (u32 k = 0; k < d; k++, row++) {
                for (u32 i = 0; i < N; i++) ha[row * N + i] = pr.uniform(q);
                for (u32 i = 0; i < N; i++)
                    he[row * N + i] = host::map_i64(q, host::gauss_round(noise));
            }
        }
    }
    for (u32 i = 0; i < N; i++) hu[i] = host::map_i64(q, h_u[i]);
    bool ok = true;
    u64 *d_a = nullptr, *d_e = nullptr, *d_prod = nullptr, *d_u = nullptr, *d_out = nullptr;
    std::vector<NiRow> rows(R);
    KG_CHK(hipMalloc(&d_a, 8 * R * N));
    KG_CHK(hipMalloc(&d_e, 8 * R * N));
    KG_CHK(hipMalloc(&d_prod, 8 * R * N));
    KG_CHK(hipMalloc(&d_out, 8 * R * N));
    KG_CHK(hipMalloc(&d_u, 8 * (size_t)N));
    KG_CHK(hipMemcpyAsync(d_a, ha.data(), 8 * R * N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_e, he.data(), 8 * R * N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_u, hu.data(), 8 * (size_t)N, hipMemcpyHostToDevice, st));
    if (!ni_products(logn, dc, st, d_a, R, h_s, N, q, d_prod)) KG_FAIL("products with the secret failed");
    for (size_t r = 0; r < R; r++) {
        NiRow &x = rows[r];
        x.prod = d_prod + r * N;
        x.e = d_e + r * N;
        x.out = d_out + r * N;
        x.msg = r < d ? d_u : nullptr;
        x.beta = r < d ? 1ull << (p.uitos_logb * r + dc.dec_uitos.ignore_bits) : 0;
        x.mono_pos = 0;
        x.mono_val = 0;
    }
    if (!ni_combine(st, dc, rows, N)) KG_FAIL("row combination failed");
    KG_CHK(pz_copy_sync(h_ksk_b, d_out, 8 * (size_t)d * N, hipMemcpyDeviceToHost, st));
    if (total > 1)
        KG_CHK(pz_copy_sync(h_zero_others, d_out + (size_t)d * N, 8 * (R - d) * N,
                            hipMemcpyDeviceToHost, st));
done:
    hipFree(d_a);
    hipFree(d_e);
    hipFree(d_prod);
    hipFree(d_out);
    hipFree(d_u);
    return ok ? PZ_OK : PZ_ERR_HIP;
}

extern "C" int pz_ni_share_rgsw(pz_ctx *c, const uint8_t *crs, const uint8_t *noise_seed,
                                uint32_t self, uint32_t total, const int32_t *h_s,
                                const int32_t *h_u, const int32_t *h_z, uint32_t l0,
                                uint32_t count, uint64_t *h_zero, uint64_t *h_cts)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!crs || !noise_seed || !h_s || !h_u || !h_z || !h_zero || !h_cts)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !crs || !noise_seed || !h_s || !h_u || !h_z || !h_zero || !h_cts");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if ((rc = ni_check(p, self, total))) return rc;
    if ((uint64_t)l0 + count > p.lwe_n)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: (uint64_t)l0 + count > p.lwe_n");
    if (count == 0) return PZ_OK;
    const u64 q = p.rlwe_q;
    const u32 N = p.rlwe_n, da = p.rgrg_da;
    const u32 dmax = p.rgrg_da > p.rgrg_db ? p.rgrg_da : p.rgrg_db;
    const u32 dmax1 = p.rlrg_da > p.rlrg_db ? p.rlrg_da : p.rlrg_db;
    const u32 lq = host::ref_log2(q), per = host::ni_per_party(p.lwe_n, total);
    // per index: a_{i,l} rows [dmax][N]; noise of the zero rows [da][N] and of
    // the ct rows [dmax][N] (rows the party does not fill stay zero)
    const size_t A = (size_t)count * dmax * N, Z = (size_t)count * da * N;
    std::vector<u64> ha(A), hez(Z), hec(A, 0);
    uint8_t nr[32];
    host::puncture(noise_seed, 2, nr);
    host::parallel_for(count, [&](size_t t) {
        const u32 l = l0 + (u32)t;
        uint8_t sd[32], nl[32];
        host::ni_seed_index(crs, l, sd);
        host::ChaCha8 pr(sd);
        for (size_t i = 0; i < (size_t)dmax * N; i++) ha[t * dmax * N + i] = pr.uniform(q);
        host::puncture(nr, l + 1, nl);
        host::ChaCha8 noise(nl);
        for (size_t i = 0; i < (size_t)da * N; i++)
            hez[t * da * N + i] = host::map_i64(q, host::gauss_round(noise));
        const u32 first = l / per == self ? dmax - dmax1 : 0;
        for (size_t i = (size_t)first * N; i < (size_t)dmax * N; i++)
            hec[t * dmax * N + i] = host::map_i64(q, host::gauss_round(noise));
    });
    bool ok = true;
    u64 *d_a = nullptr, *d_ez = nullptr, *d_ec = nullptr, *d_ps = nullptr, *d_pu = nullptr;
    u64 *d_zero = nullptr, *d_cts = nullptr;
    std::vector<NiRow> rows;
    KG_CHK(hipMalloc(&d_a, 8 * A));
    KG_CHK(hipMalloc(&d_ez, 8 * Z));
    KG_CHK(hipMalloc(&d_ec, 8 * A));
    KG_CHK(hipMalloc(&d_ps, 8 * A));
    KG_CHK(hipMalloc(&d_pu, 8 * A));
    KG_CHK(hipMalloc(&d_zero, 8 * Z));
    KG_CHK(hipMalloc(&d_cts, 8 * A));
    KG_CHK(hipMemcpyAsync(d_a, ha.data(), 8 * A, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_ez, hez.data(), 8 * Z, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_ec, hec.data(), 8 * A, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemsetAsync(d_cts, 0, 8 * A, st));
    if (!ni_products(logn, dc, st, d_a, (size_t)count * dmax, h_s, N, q, d_ps))
        KG_FAIL("products with the secret failed");
    if (!ni_products(logn, dc, st, d_a, (size_t)count * dmax, h_u, N, q, d_pu))
        KG_FAIL("products with the secret failed");
    for (u32 t = 0; t < count; t++) {
        const u32 l = l0 + t;
        const bool lead = l / per == self;
        // X^{emb * z_j[l]} (encode_x_pow_si_with_emebedding_factor, src/utils.rs:195-217)
        const int64_t si = (int64_t)dc.embedding * h_z[l];
        const u32 pos = si < 0 ? N - (u32)(-si) : (u32)si;
        const u64 val = si < 0 ? q - 1 : 1;
        for (u32 k = 0; k < da; k++) {
            const size_t ai = ((size_t)t * dmax + dmax - da + k) * N;
            rows.push_back({d_ps + ai, d_ez + ((size_t)t * da + k) * N, nullptr,
                            d_zero + ((size_t)t * da + k) * N, 0, 0, 0});
        }
        const u32 first = lead ? dmax - dmax1 : 0;
        const u32 logb = lead ? p.rlrg_logb : p.rgrg_logb;
        const u32 ib = lq - logb * (lead ? dmax1 : dmax);
        for (u32 i = first; i < dmax; i++) {
            const size_t ai = ((size_t)t * dmax + i) * N;
            rows.push_back({d_pu + ai, d_ec + ai, nullptr, d_cts + ai,
                            1ull << (ib + logb * (i - first)), pos, val});
        }
    }
    if (!ni_combine(st, dc, rows, N)) KG_FAIL("row combination failed");
    KG_CHK(pz_copy_sync(h_zero, d_zero, 8 * Z, hipMemcpyDeviceToHost, st));
    KG_CHK(pz_copy_sync(h_cts, d_cts, 8 * A, hipMemcpyDeviceToHost, st));
done:
    hipFree(d_a);
    hipFree(d_ez);
    hipFree(d_ec);
    hipFree(d_ps);
    hipFree(d_pu);
    hipFree(d_zero);
    hipFree(d_cts);
    return ok ? PZ_OK : PZ_ERR_HIP;
}


// ---- non-interactive automorphism-key and LWE-KSK shares (client side;
// _common_rountine_multi_party_auto_keys_share_gen and
// _common_rountine_multi_party_lwe_ksk_share_gen, src/bool/evaluator.rs:
// 1952-1965, 1981-2046).  The a parts are the CRS streams
// puncture(puncture(crs, 1), 2) (automorphism keys, rows in dlog order
// [-g, g, .., g^w], d_auto rows each) and puncture(puncture(crs, 1), 3) (LWE
// KSK, n entries per row, rows i*d + j); noise comes from
// ChaCha8(puncture(noise_seed, 3)) and ChaCha8(puncture(noise_seed, 4)).
extern "C" int pz_ni_share_auto_ksk(pz_ctx *c, const uint8_t *crs, const uint8_t *noise_seed,
                                    const int32_t *h_s, const int32_t *h_z, uint64_t *h_auto_b,
                                    uint64_t *h_ksk_b)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!crs || !noise_seed || !h_s || !h_z || !h_auto_b || !h_ksk_b)
        return PZ_FAIL(PZ_ERR_ARG, "null argument");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    const u64 q = p.rlwe_q;
    const u32 N = p.rlwe_n, n = p.lwe_n, dau = p.auto_d;
    const size_t R = (size_t)(p.w + 1) * dau;  // automorphism-key rows
    const size_t krows = (size_t)N * p.lwe_d;   // LWE-KSK rows
    uint8_t sd_auto[32], sd_ksk[32], n_auto[32], n_ksk[32];
    host::ni_seed_key_part(crs, 2, sd_auto);
    host::ni_seed_key_part(crs, 3, sd_ksk);
    host::puncture(noise_seed, 3, n_auto);
    host::puncture(noise_seed, 4, n_ksk);
    // automorphism keys: a rows, noise and the messages -s(X^k) on the host
    std::vector<u64> ha(R * N), he(R * N), hmsg((size_t)(p.w + 1) * N);
    {
        host::ChaCha8 pr(sd_auto), noise(n_auto);
        for (size_t i = 0; i < R * N; i++) ha[i] = pr.uniform(q);
        for (size_t i = 0; i < R * N; i++) he[i] = host::map_i64(q, host::gauss_round(noise));
    }
    for (u32 k = 0; k <= p.w; k++) {
        const uint64_t kk = dc.auto_el[k];
        for (u32 i = 0; i < N; i++) {
            uint64_t to = ((uint64_t)i * kk) % (2ull * N);
            const bool pos = to < N;
            if (!pos) to -= N;
            const u64 el = host::map_i64(q, h_s[i]);
            hmsg[(size_t)k * N + to] = pos ? (el ? q - el : 0) : el;
        }
    }
    std::vector<int64_t> herr(krows);
    {
        host::ChaCha8 noise(n_ksk);
        for (size_t r = 0; r < krows; r++)
            herr[r] = (int64_t)host::map_i64(p.lwe_q, host::gauss_round(noise));
    }
    bool ok = true;
    u64 *d_a = nullptr, *d_e = nullptr, *d_prod = nullptr, *d_msg = nullptr, *d_out = nullptr;
    u64 *d_ka = nullptr, *d_kb = nullptr;
    u32 *d_kf = nullptr;
    int32_t *d_si = nullptr, *d_zi = nullptr;
    int64_t *d_err = nullptr;
    std::vector<NiRow> rows(R);
    const u32 ib_lwe = host::ref_log2(p.lwe_q) - p.lwe_logb * p.lwe_d;
    KG_CHK(hipMalloc(&d_a, 8 * R * N));
    KG_CHK(hipMalloc(&d_e, 8 * R * N));
    KG_CHK(hipMalloc(&d_prod, 8 * R * N));
    KG_CHK(hipMalloc(&d_out, 8 * R * N));
    KG_CHK(hipMalloc(&d_msg, 8 * hmsg.size()));
    KG_CHK(hipMalloc(&d_ka, 8 * krows * n));
    KG_CHK(hipMalloc(&d_kb, 8 * krows));
    KG_CHK(hipMalloc(&d_kf, 4 * 12));
    KG_CHK(hipMalloc(&d_si, 4 * (size_t)N));
    KG_CHK(hipMalloc(&d_zi, 4 * (size_t)n));
    KG_CHK(hipMalloc(&d_err, 8 * krows));
    KG_CHK(hipMemcpyAsync(d_a, ha.data(), 8 * R * N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_e, he.data(), 8 * R * N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_msg, hmsg.data(), 8 * hmsg.size(), hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_si, h_s, 4 * (size_t)N, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_zi, h_z, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    KG_CHK(hipMemcpyAsync(d_err, herr.data(), 8 * krows, hipMemcpyHostToDevice, st));
    // b = a*s + e + beta_j * (-s(X^k))  (seeded_rlwe_ksk_gen, src/rgsw/keygen.rs:306-358)
    if (!ni_products(logn, dc, st, d_a, R, h_s, N, q, d_prod)) KG_FAIL("products with the secret failed");
    for (size_t r = 0; r < R; r++) {
        const u32 k = (u32)(r / dau), j = (u32)(r % dau);
        rows[r] = {d_prod + r * N, d_e + r * N, d_msg + (size_t)k * N, d_out + r * N,
                   1ull << (p.auto_logb * j + dc.dec_auto.ignore_bits), 0, 0};
    }
    if (!ni_combine(st, dc, rows, N)) KG_FAIL("row combination failed");
    // LWE KSK: the a rows expanded from the CRS stream on the device
    if (!pz_internal_expand_uniform(st, sd_ksk, 1, krows * n, 0, p.lwe_q, d_ka, d_kf))
        KG_FAIL("expansion of the KSK a parts failed");
    hipLaunchKernelGGL(k_ksk_share_b, (int)std::min<size_t>(krows, 65535), 64, 0, st, d_ka, d_si,
                       d_zi, d_err, N, p.lwe_d, n, p.lwe_logb, ib_lwe, (u64)(p.lwe_q - 1), d_kb);
    KG_CHK(hipGetLastError());
    KG_CHK(pz_copy_sync(h_auto_b, d_out, 8 * R * N, hipMemcpyDeviceToHost, st));
    KG_CHK(pz_copy_sync(h_ksk_b, d_kb, 8 * krows, hipMemcpyDeviceToHost, st));
done:
    hipFree(d_a);
    hipFree(d_e);
    hipFree(d_prod);
    hipFree(d_out);
    hipFree(d_msg);
    hipFree(d_ka);
    hipFree(d_kb);
    hipFree(d_kf);
    hipFree(d_si);
    hipFree(d_zi);
    hipFree(d_err);
    return ok ? PZ_OK : PZ_ERR_HIP;  // message set by KG_CHK / KG_FAIL
}
