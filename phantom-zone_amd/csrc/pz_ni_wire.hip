// pz_ni_wire.hip -- seeded wire formats of the non-interactive multi-party
// server key: a party's share and the aggregated, seeded server key.  Both
// carry b parts only; every a part is regenerated from the common reference
// seed, as in the reference's
//   CommonReferenceSeededNonInteractiveMultiPartyServerKeyShare
//   (src/bool/keys.rs:1240-1394) and
//   SeededNonInteractiveMultiPartyServerKey (src/bool/keys.rs:1029-1063).
// A blob is a pz_ni_blob_header followed by its sections, arrays of
// little-endian u64 words at 8-byte aligned offsets (the section table in the
// header gives each offset and length).  The share generator and the
// aggregation run on the device through the library's other entry points;
// this file packs, checks and unpacks.
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/pz.h"
#include "pz_host.hpp"
#include "pz_internal.hpp"

using namespace pz;

extern int pz_internal_ctx_view(pz_ctx *c, pz_params *p, DevCtx *dc, int *logn, hipStream_t *s);

static_assert(sizeof(pz_params) == 88, "pz_params is part of the wire format");
static_assert(sizeof(pz_ni_blob_header) == 272, "header layout");

namespace {

const char SHARE_MAGIC[8] = {'P', 'Z', 'N', 'I', 'S', 'H', 'R', 0};
const char KEY_MAGIC[8] = {'P', 'Z', 'N', 'I', 'K', 'E', 'Y', 0};

bool ni_params(const pz_params &p)
{
    return p.uitos_d && p.rgrg_da && p.rgrg_db && p.rlwe_n && p.lwe_n;
}

uint32_t dmax_rg(const pz_params &p) { return std::max(p.rgrg_da, p.rgrg_db); }
uint32_t dmax_rl(const pz_params &p) { return std::max(p.rlrg_da, p.rlrg_db); }

// [begin, end) of the LWE indices party `self` leads
// (multi_party_user_id_lwe_segment, src/bool/evaluator.rs:613-626)
void segment(const pz_params &p, uint32_t self, uint32_t total, uint32_t *b, uint32_t *e)
{
    const uint32_t per = host::ni_per_party(p.lwe_n, total);
    *b = std::min(per * self, p.lwe_n);
    *e = std::min(per * (self + 1), p.lwe_n);
}

// section lengths in u64 words, in wire order
void share_sections(const pz_params &p, uint32_t self, uint32_t total, uint64_t *w)
{
    const uint64_t N = p.rlwe_n;
    uint32_t b, e;
    segment(p, self, total, &b, &e);
    w[0] = (uint64_t)(e - b) * dmax_rl(p) * N;              // self_leader_ni_rgsw_cts
    w[1] = (uint64_t)(p.lwe_n - (e - b)) * dmax_rg(p) * N;  // not_self_leader_ni_rgsw_cts
    w[2] = (uint64_t)p.lwe_n * p.rgrg_da * N;               // ni_rgsw_zero_encs
    w[3] = (uint64_t)p.uitos_d * N;                         // ui_to_s_ksk (b parts)
    w[4] = (uint64_t)(total - 1) * p.uitos_d * N;           // ksk_zero_encs_for_others
    w[5] = (uint64_t)(p.w + 1) * p.auto_d * N;              // auto_keys_share
    w[6] = (uint64_t)N * p.lwe_d;                           // lwe_ksk_share
}
const uint32_t SHARE_SECTIONS = 7;

void key_sections(const pz_params &p, uint32_t total, uint64_t *w)
{
    const uint64_t N = p.rlwe_n;
    w[0] = (uint64_t)total * p.uitos_d * N;                               // ui_to_s_ksks
    w[1] = (uint64_t)p.lwe_n * (2 * p.rlrg_da + 2 * p.rlrg_db) * N;      // rgsw_cts
    w[2] = (uint64_t)(p.w + 1) * p.auto_d * N;                           // auto_keys
    w[3] = (uint64_t)N * p.lwe_d;                                        // lwe_ksk
}
const uint32_t KEY_SECTIONS = 4;

size_t layout(pz_ni_blob_header *h, const uint64_t *w, uint32_t ns)
{
    size_t off = sizeof(pz_ni_blob_header);
    for (uint32_t i = 0; i < PZ_NI_SECTIONS; i++) {
        h->offset[i] = i < ns ? off : 0;
        h->words[i] = i < ns ? w[i] : 0;
        if (i < ns) off += 8 * w[i];
    }
    h->nsections = ns;
    return off;
}

uint64_t *section(uint8_t *blob, const pz_ni_blob_header &h, uint32_t i)
{
    return reinterpret_cast<uint64_t *>(blob + h.offset[i]);
}
const uint64_t *section(const uint8_t *blob, const pz_ni_blob_header &h, uint32_t i)
{
    return reinterpret_cast<const uint64_t *>(blob + h.offset[i]);
}

// header and section table of a blob against what the parameter set implies
int check_blob(const uint8_t *blob, const char *magic, const pz_params &p, pz_ni_blob_header *h,
               const char *fn)
{
    if (!blob || ((uintptr_t)blob & 7))
        return pz_internal_fail(PZ_ERR_ARG, fn, "blob is null or not 8-byte aligned");
    memcpy(h, blob, sizeof(*h));
    if (memcmp(h->magic, magic, 8) != 0)
        return pz_internal_fail(PZ_ERR_ARG, fn, "not a blob of this kind (magic)");
    if (h->version != PZ_NI_BLOB_VERSION)
        return pz_internal_fail(PZ_ERR_UNSUPPORTED, fn, "blob version not supported");
    if (memcmp(&h->params, &p, sizeof(p)) != 0)
        return pz_internal_fail(PZ_ERR_ARG, fn, "blob made for another parameter set");
    if (h->total == 0 || h->total > PZ_NI_MAX_PARTIES || h->user >= h->total)
        return pz_internal_fail(PZ_ERR_ARG, fn, "party count or user id out of range");
    uint64_t w[PZ_NI_SECTIONS] = {};
    uint32_t ns;
    if (memcmp(magic, SHARE_MAGIC, 8) == 0) {
        share_sections(p, h->user, h->total, w);
        ns = SHARE_SECTIONS;
    } else {
        key_sections(p, h->total, w);
        ns = KEY_SECTIONS;
    }
    pz_ni_blob_header want;
    memset(&want, 0, sizeof(want));
    layout(&want, w, ns);
    if (h->nsections != ns || memcmp(h->offset, want.offset, sizeof(want.offset)) != 0 ||
        memcmp(h->words, want.words, sizeof(want.words)) != 0)
        return pz_internal_fail(PZ_ERR_ARG, fn, "section table does not match the parameter set");
    return PZ_OK;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { hipFree(p); }
    bool alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 8) == hipSuccess; }
    u64 *u() const { return (u64 *)p; }
};

}  // namespace

extern "C" int pz_ni_share_bytes(const pz_params *p, uint32_t self, uint32_t total,
                                 size_t *bytes)
{
    if (!p || !bytes) return PZ_FAIL(PZ_ERR_ARG, "null argument");
    if (!ni_params(*p)) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "not a non-interactive parameter set");
    if (total == 0 || total > PZ_NI_MAX_PARTIES || self >= total)
        return PZ_FAIL(PZ_ERR_ARG, "party count or user id out of range");
    pz_ni_blob_header h;
    uint64_t w[PZ_NI_SECTIONS] = {};
    share_sections(*p, self, total, w);
    *bytes = layout(&h, w, SHARE_SECTIONS);
    return PZ_OK;
}

extern "C" int pz_ni_seeded_key_bytes(const pz_params *p, uint32_t total, size_t *bytes)
{
    if (!p || !bytes) return PZ_FAIL(PZ_ERR_ARG, "null argument");
    if (!ni_params(*p)) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "not a non-interactive parameter set");
    if (total == 0 || total > PZ_NI_MAX_PARTIES)
        return PZ_FAIL(PZ_ERR_ARG, "party count out of range");
    pz_ni_blob_header h;
    uint64_t w[PZ_NI_SECTIONS] = {};
    key_sections(*p, total, w);
    *bytes = layout(&h, w, KEY_SECTIONS);
    return PZ_OK;
}

// gen_non_interactive_multi_party_key_share (src/bool/evaluator.rs:1624-1979)
extern "C" int pz_ni_key_share(pz_ctx *c, const uint8_t *crs, const uint8_t *noise_seed,
                               uint32_t self, uint32_t total, const int32_t *h_s,
                               const int32_t *h_u, const int32_t *h_z, uint8_t *h_share)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!crs || !noise_seed || !h_s || !h_u || !h_z || !h_share)
        return PZ_FAIL(PZ_ERR_ARG, "null argument");
    if ((uintptr_t)h_share & 7) return PZ_FAIL(PZ_ERR_ARG, "h_share must be 8-byte aligned");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (!ni_params(p)) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "not a non-interactive parameter set");
    if (total == 0 || total > PZ_NI_MAX_PARTIES || self >= total)
        return PZ_FAIL(PZ_ERR_ARG, "party count or user id out of range");
    pz_ni_blob_header h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, SHARE_MAGIC, 8);
    h.version = PZ_NI_BLOB_VERSION;
    h.user = self;
    h.total = total;
    h.params = p;
    memcpy(h.crs, crs, 32);
    uint64_t w[PZ_NI_SECTIONS] = {};
    share_sections(p, self, total, w);
    layout(&h, w, SHARE_SECTIONS);
    memcpy(h_share, &h, sizeof(h));
    const uint32_t N = p.rlwe_n, n = p.lwe_n, dm = dmax_rg(p), d1 = dmax_rl(p);
    // u -> s parts
    std::vector<uint64_t> zero_others(total > 1 ? w[4] : (uint64_t)p.uitos_d * N);
    rc = pz_ni_share_uitos(c, crs, noise_seed, self, total, h_s, h_u, section(h_share, h, 3),
                           zero_others.data());
    if (rc) return rc;
    if (total > 1) memcpy(section(h_share, h, 4), zero_others.data(), 8 * w[4]);
    // RGSW parts for every index, then packed: the rows a leader fills
    // (the last max(rlrg_da, rlrg_db) of the index's stream) and all rows
    // elsewhere, in ascending index order
    std::vector<uint64_t> cts((size_t)n * dm * N);
    rc = pz_ni_share_rgsw(c, crs, noise_seed, self, total, h_s, h_u, h_z, 0, n,
                          section(h_share, h, 2), cts.data());
    if (rc) return rc;
    uint32_t b, e;
    segment(p, self, total, &b, &e);
    uint64_t *lead = section(h_share, h, 0), *rest = section(h_share, h, 1);
    for (uint32_t l = 0; l < n; l++) {
        const uint64_t *ct = cts.data() + (size_t)l * dm * N;
        if (l >= b && l < e) {
            memcpy(lead, ct + (size_t)(dm - d1) * N, 8 * (size_t)d1 * N);
            lead += (size_t)d1 * N;
        } else {
            memcpy(rest, ct, 8 * (size_t)dm * N);
            rest += (size_t)dm * N;
        }
    }
    // automorphism keys and LWE KSK
    return pz_ni_share_auto_ksk(c, crs, noise_seed, h_s, h_z, section(h_share, h, 5),
                                section(h_share, h, 6));
}

// aggregate_non_interactive_multi_party_server_key_shares
// (src/bool/evaluator.rs:1222-1622)
extern "C" int pz_ni_aggregate_shares(pz_ctx *c, const uint8_t *const *h_shares, uint32_t total,
                                      uint8_t *h_key)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!h_shares || !h_key) return PZ_FAIL(PZ_ERR_ARG, "null argument");
    if ((uintptr_t)h_key & 7) return PZ_FAIL(PZ_ERR_ARG, "h_key must be 8-byte aligned");
    if (total == 0 || total > PZ_NI_MAX_PARTIES)
        return PZ_FAIL(PZ_ERR_ARG, "party count out of range");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (!ni_params(p)) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "not a non-interactive parameter set");
    // every share checked, then indexed by its user id
    std::vector<pz_ni_blob_header> hs(total);
    std::vector<const uint8_t *> by_user(total, nullptr);
    for (uint32_t t = 0; t < total; t++) {
        if ((rc = check_blob(h_shares[t], SHARE_MAGIC, p, &hs[t], __func__))) return rc;
        if (hs[t].total != total)
            return PZ_FAIL(PZ_ERR_ARG, "a share was made for another number of parties");
        if (memcmp(hs[t].crs, hs[0].crs, 32) != 0)
            return PZ_FAIL(PZ_ERR_ARG, "shares of different common reference seeds");
        if (by_user[hs[t].user]) return PZ_FAIL(PZ_ERR_ARG, "two shares of the same party");
        by_user[hs[t].user] = h_shares[t];
    }
    const uint8_t *crs = hs[0].crs;
    pz_ni_blob_header sh[PZ_NI_MAX_PARTIES];
    for (uint32_t j = 0; j < total; j++) memcpy(&sh[j], by_user[j], sizeof(sh[j]));
    const uint32_t N = p.rlwe_n, n = p.lwe_n, dm = dmax_rg(p), d1 = dmax_rl(p);
    const size_t uw = (size_t)p.uitos_d * N;
    pz_ni_blob_header kh;
    memset(&kh, 0, sizeof(kh));
    memcpy(kh.magic, KEY_MAGIC, 8);
    kh.version = PZ_NI_BLOB_VERSION;
    kh.total = total;
    kh.params = p;
    memcpy(kh.crs, crs, 32);
    uint64_t kw[PZ_NI_SECTIONS] = {};
    key_sections(p, total, kw);
    layout(&kh, kw, KEY_SECTIONS);
    memcpy(h_key, &kh, sizeof(kh));
    // u_j -> s key-switching keys: user j's ksk_b plus the others' zero
    // encryptions for j (src/bool/evaluator.rs:1266-1337); the seeded key
    // keeps their b parts in the coefficient domain (:1563-1569)
    std::vector<pz_rlwe_ksk *> ksks(total, nullptr);
    auto free_ksks = [&]() {
        for (auto *k : ksks) pz_rlwe_ksk_free(k);
    };
    {
        DevBuf d_b, d_z;
        if (!d_b.alloc(8 * uw) || !d_z.alloc(8 * uw * std::max<uint32_t>(total - 1, 1)))
            return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
        uint64_t *kb = section(h_key, kh, 0);
        for (uint32_t j = 0; j < total; j++) {
            const uint64_t *own = section(by_user[j], sh[j], 3);
            uint64_t *bj = kb + (size_t)j * uw;
            memcpy(bj, own, 8 * uw);
            std::vector<uint64_t> zf;
            for (uint32_t o = 0; o < total; o++) {
                if (o == j) continue;
                const uint64_t *z = section(by_user[o], sh[o], 4) + (size_t)(j < o ? j : j - 1) * uw;
                zf.insert(zf.end(), z, z + uw);
                for (size_t i = 0; i < uw; i++) {
                    const uint64_t v = bj[i] + z[i];
                    bj[i] = v >= p.rlwe_q ? v - p.rlwe_q : v;
                }
            }
            if (pz_copy_sync(d_b.p, own, 8 * uw, hipMemcpyHostToDevice, st) != hipSuccess ||
                (total > 1 &&
                 pz_copy_sync(d_z.p, zf.data(), 8 * zf.size(), hipMemcpyHostToDevice, st) != hipSuccess)) {
                free_ksks();
                return PZ_FAIL(PZ_ERR_HIP, "copy failed");
            }
            if ((rc = pz_ni_uitos_ksk(c, crs, total, j, d_b.u(), total > 1 ? d_z.u() : nullptr,
                                      &ksks[j]))) {
                free_ksks();
                return rc;
            }
        }
    }
    // RGSWs: each party's zero encryptions and NI RGSW rows (a leader's
    // packed rows back in the last rows of the index's stream)
    {
        const size_t zw = (size_t)n * p.rgrg_da * N, cw = (size_t)n * dm * N;
        std::vector<uint64_t> cts(cw);
        DevBuf d_zero, d_cts, d_out;
        if (!d_zero.alloc(8 * zw * total) || !d_cts.alloc(8 * cw * total) ||
            !d_out.alloc(8 * kw[1])) {
            free_ksks();
            return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
        }
        bool ok = true;
        for (uint32_t j = 0; j < total && ok; j++) {
            uint32_t b, e;
            segment(p, j, total, &b, &e);
            const uint64_t *lead = section(by_user[j], sh[j], 0);
            const uint64_t *rest = section(by_user[j], sh[j], 1);
            for (uint32_t l = 0; l < n; l++) {
                uint64_t *ct = cts.data() + (size_t)l * dm * N;
                if (l >= b && l < e) {
                    memset(ct, 0, 8 * (size_t)(dm - d1) * N);
                    memcpy(ct + (size_t)(dm - d1) * N, lead, 8 * (size_t)d1 * N);
                    lead += (size_t)d1 * N;
                } else {
                    memcpy(ct, rest, 8 * (size_t)dm * N);
                    rest += (size_t)dm * N;
                }
            }
            ok = pz_copy_sync(d_zero.u() + j * zw, section(by_user[j], sh[j], 2), 8 * zw,
                           hipMemcpyHostToDevice, st) == hipSuccess &&
                 pz_copy_sync(d_cts.u() + j * cw, cts.data(), 8 * cw, hipMemcpyHostToDevice, st) ==
                     hipSuccess;
        }
        if (!ok) {
            free_ksks();
            return PZ_FAIL(PZ_ERR_HIP, "copy failed");
        }
        rc = pz_ni_aggregate_rgsw(c, crs, total, ksks.data(), 0, n, d_zero.u(), d_cts.u(),
                                  d_out.u(), nullptr);
        free_ksks();
        if (rc) return rc;
        if (pz_copy_sync(section(h_key, kh, 1), d_out.p, 8 * kw[1], hipMemcpyDeviceToHost, st) !=
            hipSuccess)
            return PZ_FAIL(PZ_ERR_HIP, "copy failed");
    }
    // automorphism keys and LWE KSK: sums of the b parts
    {
        DevBuf d_as, d_ks, d_ab, d_kb;
        if (!d_as.alloc(8 * kw[2] * total) || !d_ks.alloc(8 * kw[3] * total) ||
            !d_ab.alloc(8 * kw[2]) || !d_kb.alloc(8 * kw[3]))
            return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
        for (uint32_t j = 0; j < total; j++)
            if (pz_copy_sync(d_as.u() + j * kw[2], section(by_user[j], sh[j], 5), 8 * kw[2],
                          hipMemcpyHostToDevice, st) != hipSuccess ||
                pz_copy_sync(d_ks.u() + j * kw[3], section(by_user[j], sh[j], 6), 8 * kw[3],
                          hipMemcpyHostToDevice, st) != hipSuccess)
                return PZ_FAIL(PZ_ERR_HIP, "copy failed");
        if ((rc = pz_ni_aggregate_auto_ksk(c, total, d_as.u(), d_ks.u(), d_ab.u(), d_kb.u(),
                                           nullptr)))
            return rc;
        if (pz_copy_sync(section(h_key, kh, 2), d_ab.p, 8 * kw[2], hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            pz_copy_sync(section(h_key, kh, 3), d_kb.p, 8 * kw[3], hipMemcpyDeviceToHost, st) !=
                hipSuccess)
            return PZ_FAIL(PZ_ERR_HIP, "copy failed");
    }
    return PZ_OK;
}

// NonInteractiveServerKeyEvaluationDomain::from (src/bool/keys.rs:851-1005)
// and the Shoup twins (:1088-1155), from a seeded server key blob
extern "C" int pz_ni_server_key_from_seeded(pz_ctx *c, const uint8_t *h_key, pz_key **key,
                                            pz_rlwe_ksk **ksks)
{
    pz_params p;
    DevCtx dc;
    int logn;
    hipStream_t st;
    if (!h_key || !key) return PZ_FAIL(PZ_ERR_ARG, "null argument");
    int rc = pz_internal_ctx_view(c, &p, &dc, &logn, &st);
    if (rc) return rc;
    if (!ni_params(p)) return PZ_FAIL(PZ_ERR_UNSUPPORTED, "not a non-interactive parameter set");
    pz_ni_blob_header kh;
    if ((rc = check_blob(h_key, KEY_MAGIC, p, &kh, __func__))) return rc;
    {
        DevBuf d_rg, d_ab, d_kb;
        if (!d_rg.alloc(8 * kh.words[1]) || !d_ab.alloc(8 * kh.words[2]) ||
            !d_kb.alloc(8 * kh.words[3]))
            return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
        if (pz_copy_sync(d_rg.p, section(h_key, kh, 1), 8 * kh.words[1], hipMemcpyHostToDevice, st) !=
                hipSuccess ||
            pz_copy_sync(d_ab.p, section(h_key, kh, 2), 8 * kh.words[2], hipMemcpyHostToDevice, st) !=
                hipSuccess ||
            pz_copy_sync(d_kb.p, section(h_key, kh, 3), 8 * kh.words[3], hipMemcpyHostToDevice, st) !=
                hipSuccess)
            return PZ_FAIL(PZ_ERR_HIP, "copy failed");
        if ((rc = pz_ni_expand_server_key(c, kh.crs, d_rg.u(), d_ab.u(), d_kb.u(), key)))
            return rc;
    }
    if (!ksks) return PZ_OK;
    // u_j -> s keys: A rows from user j's seed, B rows the stored b parts
    // (keys.rs:947-992); the aggregation already added the zero encryptions,
    // so the "others" passed here are zero rows
    const size_t uw = (size_t)p.uitos_d * p.rlwe_n;
    DevBuf d_b, d_z;
    if (!d_b.alloc(8 * uw) || !d_z.alloc(8 * uw * std::max<uint32_t>(kh.total - 1, 1)) ||
        hipMemsetAsync(d_z.p, 0, 8 * uw * std::max<uint32_t>(kh.total - 1, 1), st) !=
            hipSuccess) {
        pz_key_free(*key);
        *key = nullptr;
        return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
    }
    for (uint32_t j = 0; j < kh.total; j++) ksks[j] = nullptr;
    for (uint32_t j = 0; j < kh.total; j++) {
        rc = pz_copy_sync(d_b.p, section(h_key, kh, 0) + (size_t)j * uw, 8 * uw,
                       hipMemcpyHostToDevice, st) == hipSuccess
                 ? pz_ni_uitos_ksk(c, kh.crs, kh.total, j, d_b.u(),
                                   kh.total > 1 ? d_z.u() : nullptr, &ksks[j])
                 : PZ_FAIL(PZ_ERR_HIP, "copy failed");
        if (rc) {
            for (uint32_t o = 0; o < j; o++) pz_rlwe_ksk_free(ksks[o]);
            pz_key_free(*key);
            *key = nullptr;
            return rc;
        }
    }
    return PZ_OK;
}
