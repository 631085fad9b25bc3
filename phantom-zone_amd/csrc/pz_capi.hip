// pz_capi.hip -- the extern "C" boundary (include/pz.h): contexts, resident
// server keys, and the launch sequences of every operator.
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pz.h"
#include "pz_host.hpp"
#include "pz_internal.hpp"
#include "pz_exec.hpp"

using namespace pz;

static thread_local std::string g_err;

// Records the message pz_last_error() returns: "<function>: <what>", plus the
// HIP runtime's text for PZ_ERR_HIP when a HIP call left an error behind.
int pz_internal_fail(int code, const char *fn, const char *what)
{
    g_err = std::string(fn) + ": " + what;
    if (code == PZ_ERR_HIP) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) g_err += std::string(" (") + hipGetErrorString(e) + ")";
    }
    return code;
}

int pz_internal_launch_status(const char *fn)
{
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return PZ_OK;
    g_err = std::string(fn) + ": kernel launch failed (" + hipGetErrorString(e) + ")";
    return PZ_ERR_HIP;
}

#define fail(code, msg) pz_internal_fail((code), __func__, (msg))

#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            g_err = std::string(__func__) + ": " #x ": " + hipGetErrorString(e_);      \
            return PZ_ERR_HIP;                                                        \
        }                                                                             \
    } while (0)

struct GraphKey {
    const void *ops, *c0, *c1, *out;
    size_t B;
    bool timed;  // captured with the per-kernel timing events
    bool operator<(const GraphKey &o) const
    {
        return std::tie(ops, c0, c1, out, B, timed) <
               std::tie(o.ops, o.c0, o.c1, o.out, o.B, o.timed);
    }
};

struct pz_ctx {
    pz_params p;
    int device = 0;
    int logn = 0;
    hipStream_t stream = nullptr;
    DevCtx dc;
    host::HostTables ht;
    // device allocations of the tables
    u64 *d_tables = nullptr;   // 8N
    u32 *d_dlog = nullptr;
    u64 *d_tv = nullptr;
    // scratch of the PBS pipeline
    int8_t *d_digits = nullptr;
    u32 *d_ksb = nullptr, *d_ks = nullptr, *d_stats = nullptr;
    size_t cap = 0;
    int num_cu = 0;
    int br_per_cu[BR_NVAR] = {};  // resident work groups per CU of each kernel variant
                                  // (BR_ONE, BR_SPLIT_LAT, BR_SPLIT, BR_SPLIT4_LAT,
                                  // BR_SPLIT4; 0: not available)
    int br_mode = 0;  // 0 auto, 1 one-half kernel, 2 split kernel, 3 split with 4 coefficients
                      // per thread
    // autotuned blind-rotation shapes per batch bucket (ceil(log2 B)): kernel
    // variant and resident work groups per CU; var < 0 = not tuned
    struct Tuned {
        int var = -1, per_cu = 0;
    } tuned[40];
    // key-index order of a batched RLWE x RGSW / automorphism launch (B
    // words), grown on demand
    u32 *d_order = nullptr;
    size_t order_cap = 0;
    // workspace of the key-preparation operators (grow-only; every user
    // synchronises the stream before it returns)
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // captured gate batches, all made with the key whose id is graph_key_id
    // (ids are never reused, so a key freed elsewhere -- its graphs then hold
    // dangling pointers -- can never match a later key at the same address)
    std::map<GraphKey, hipGraphExec_t> graphs;
    uint64_t graph_key_id = 0;
    // timing
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float t_prep = 0, t_ks = 0, t_br = 0;
};

static std::atomic<uint64_t> g_next_key_id{1};

struct pz_key {
    pz_ctx *owner = nullptr;
    uint64_t id = g_next_key_id.fetch_add(1);  // unique for the process's lifetime
    DevKey dk{};
    u64 *rgsw = nullptr, *rgsw_s = nullptr, *auto_k = nullptr, *auto_s = nullptr;
    void *ksk = nullptr;
    int8_t *ksk_planes = nullptr;
    size_t bytes = 0;
};

extern "C" const char *pz_last_error(void) { return g_err.c_str(); }

bool pz_internal_key_fits(const pz_ctx *c, const pz_key *k);
#define KEY_OK(c, k)                                                                \
    do {                                                                            \
        if (!pz_internal_key_fits(c, k))                                            \
            return fail(PZ_ERR_ARG, "key made for another parameter set or device"); \
    } while (0)

// Pipeline scratch of one gate / PBS batch: key-switch digits (tile-major),
// bodies, key-switch outputs and automorphism counts.
struct PbsScratch {
    int8_t *digits;
    u32 *ksb, *ks, *stats;
};

// ---- fixed pipeline scratch of pz_gates_noalloc: one array per GPU, held
// until the process ends.  Host code never passes the symbol itself (in host
// code it names a host-side stand-in): each device's address comes from
// hipGetSymbolAddress at the first context creation on that device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Records the device's scratch address; `device` must be current.
static int noalloc_setup(int device)
{
    if (device >= NOALLOC_MAX_DEVICES) return PZ_OK;  // pz_gates_noalloc refuses it
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return PZ_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(PZ_ERR_HIP, "address of the fixed scratch not available");
    g_scratch_ptr[device] = ptr;
    return PZ_OK;
}

// Carves the pipeline scratch of `cap` ciphertexts (a multiple of the key
// switch's tile) out of `base`, each part 256-byte aligned; returns the bytes
// used (base = nullptr only measures).
static size_t noalloc_layout(const pz_params &p, size_t cap, uint8_t *base, PbsScratch *s)
{
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t digits = up((size_t)p.rlwe_n * p.lwe_d * cap), ksb = up(4 * cap);
    const size_t ks = up(4 * cap * ((size_t)p.lwe_n + 1)), stats = up(4 * cap);
    if (s && base) {
        s->digits = (int8_t *)base;
        s->ksb = (u32 *)(base + digits);
        s->ks = (u32 *)(base + digits + ksb);
        s->stats = (u32 *)(base + digits + ksb + ks);
    }
    return digits + ksb + ks + stats;
}

// Largest batch the fixed scratch holds for a parameter set.
static size_t noalloc_max_batch(const pz_params &p)
{
    size_t cap = 0;
    while (noalloc_layout(p, cap + KS_CT, nullptr, nullptr) <= NOALLOC_SCRATCH_BYTES) cap += KS_CT;
    return cap;
}

extern "C" int pz_params_named(const char *name, pz_params *out)
{
    if (!name || !out) return fail(PZ_ERR_ARG, "null argument");
    if (host::params_named(name, out) != 0) return fail(PZ_ERR_ARG, "unknown parameter set");
    return PZ_OK;
}

static int check_params(const pz_params *p, int *logn)
{
    uint32_t N = p->rlwe_n;
    int l = 0;
    while ((1u << l) < N) l++;
    if ((1u << l) != N || (l != 9 && l != 11))
        return fail(PZ_ERR_UNSUPPORTED, "rlwe_n must be 512 or 2048");
    // unreduced transform stages keep values below 2^10 q (inverse) and 23 q
    // (forward), so q < 2^54
    if (p->rlwe_q >= (1ull << 54) || p->rlwe_q % (2ull * N) != 1)
        return fail(PZ_ERR_ARG, "rlwe_q must be < 2^54 and 1 mod 2N");
    // the N = 2048 transforms reduce with floor(2^64/q) as a 32-bit factor
    if (l == 11 && p->rlwe_q <= (1ull << 32))
        return fail(PZ_ERR_UNSUPPORTED, "rlwe_q must exceed 2^32 at N = 2048");
    if (p->lwe_q & (p->lwe_q - 1) || p->lwe_q > (1ull << 17) || p->lwe_q < 2)
        return fail(PZ_ERR_UNSUPPORTED, "lwe_q must be a power of two <= 2^17");
    if (p->br_q & (p->br_q - 1) || p->br_q > 2 * N || p->br_q < 16)
        return fail(PZ_ERR_ARG, "br_q must be a power of two <= 2N");
    if (p->w + 1 > (uint32_t)MAX_AUTO) return fail(PZ_ERR_UNSUPPORTED, "window too large");
    if (p->lwe_n == 0 || p->lwe_n > 4096) return fail(PZ_ERR_UNSUPPORTED, "lwe_n out of range");
    if (p->lwe_logb > 7 || p->lwe_logb == 0) return fail(PZ_ERR_UNSUPPORTED, "lwe_logb must be 1..7");
    if (p->rlrg_da == 0 || p->rlrg_db == 0 || p->auto_d == 0 || p->lwe_d == 0)
        return fail(PZ_ERR_ARG, "decomposition counts must be > 0");
    // digits are formed in 32-bit lanes (a digit and its carry bit within 31 bits)
    if (p->rlrg_logb == 0 || p->rlrg_logb > 30 || p->auto_logb == 0 || p->auto_logb > 30 ||
        p->uitos_logb > 30 || p->rgrg_logb > 30)
        return fail(PZ_ERR_UNSUPPORTED, "decomposition bases must be 2^1 .. 2^30");
    uint32_t lq = host::ref_log2(p->rlwe_q);
    if (p->rlrg_logb * p->rlrg_da > lq || p->rlrg_logb * p->rlrg_db > lq ||
        p->auto_logb * p->auto_d > lq || p->lwe_logb * p->lwe_d > host::ref_log2(p->lwe_q))
        return fail(PZ_ERR_ARG, "decomposer wants logq >= logb*d");
    if (p->uitos_d && (p->uitos_logb == 0 || p->uitos_logb * p->uitos_d > lq || p->uitos_d > 64))
        return fail(PZ_ERR_ARG, "u -> s decomposer wants 0 < logb*d <= logq, d <= 64");
    if ((p->rgrg_da != 0) != (p->rgrg_db != 0) ||
        (p->rgrg_da && (p->rgrg_logb == 0 || p->rgrg_logb * p->rgrg_da > lq ||
                        p->rgrg_logb * p->rgrg_db > lq)))
        return fail(PZ_ERR_ARG, "RGSW x RGSW decomposer wants 0 < logb*d <= logq for both counts");
    *logn = l;
    return PZ_OK;
}

static Decomp make_decomp(uint64_t q, uint32_t logb, uint32_t d)
{
    Decomp dc;
    dc.q = q;
    dc.logb = logb;
    dc.d = d;
    dc.ignore_bits = host::ref_log2(q) - logb * d;
    decomp_centring(dc);
    return dc;
}

extern "C" int pz_ctx_create(const pz_params *p, int device, pz_ctx **out)
{
    if (!p || !out) return fail(PZ_ERR_ARG, "null argument");
    int logn = 0;
    int rc = check_params(p, &logn);
    if (rc) return rc;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(PZ_ERR_ARG, "no such device");
    HIPCHK(hipSetDevice(device));
    if ((rc = noalloc_setup(device))) return rc;
    pz_ctx *c = new pz_ctx();
    c->p = *p;
    c->device = device;
    c->logn = logn;
    uint32_t N = p->rlwe_n;
    c->ht = host::make_tables(p->rlwe_q, N);
    if (c->ht.psi == 0) {
        delete c;
        return fail(PZ_ERR_ARG, "no 2N-th root of unity found");
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(PZ_ERR_HIP, "stream create failed");
    }
    // [fw pairs (2N) | iw pairs (2N) | fw63, iw63 pairs (2N each): the powers
    // with floor(w 2^63 / q), the twins of FWD_FAST6 and of the split step's
    // fast inverse]
    std::vector<u64> tabs(8 * (size_t)N);
    for (uint32_t i = 0; i < N; i++) {
        tabs[2 * i] = c->ht.psi_bo[i];
        tabs[2 * i + 1] = c->ht.psi_bo_s[i];
        tabs[2 * N + 2 * i] = c->ht.ipsi_bo[i];
        tabs[2 * N + 2 * i + 1] = c->ht.ipsi_bo_s[i];
        tabs[4 * N + 2 * i] = c->ht.psi_bo[i];
        tabs[4 * N + 2 * i + 1] =
            (u64)(((unsigned __int128)c->ht.psi_bo[i] << 63) / p->rlwe_q);
        tabs[6 * N + 2 * i] = c->ht.ipsi_bo[i];
        tabs[6 * N + 2 * i + 1] =
            (u64)(((unsigned __int128)c->ht.ipsi_bo[i] << 63) / p->rlwe_q);
    }
    std::vector<uint32_t> dlog = host::dlog_map(*p);
    std::vector<uint64_t> tv = host::test_vectors(*p);
    if (hipMalloc(&c->d_tables, 8 * tabs.size()) != hipSuccess ||
        hipMalloc(&c->d_dlog, 4 * dlog.size()) != hipSuccess ||
        hipMalloc(&c->d_tv, 8 * tv.size()) != hipSuccess) {
        pz_ctx_destroy(c);
        return fail(PZ_ERR_HIP, "table allocation failed");
    }
    hipMemcpy(c->d_tables, tabs.data(), 8 * tabs.size(), hipMemcpyHostToDevice);
    hipMemcpy(c->d_dlog, dlog.data(), 4 * dlog.size(), hipMemcpyHostToDevice);
    hipMemcpy(c->d_tv, tv.data(), 8 * tv.size(), hipMemcpyHostToDevice);

    DevCtx &d = c->dc;
    memset(&d, 0, sizeof(d));
    d.tb.fw = reinterpret_cast<const u64x2 *>(c->d_tables);
    d.tb.iw = reinterpret_cast<const u64x2 *>(c->d_tables + 2 * N);
    d.tb.fw63 = reinterpret_cast<const u64x2 *>(c->d_tables + 4 * N);
    d.tb.iw63 = reinterpret_cast<const u64x2 *>(c->d_tables + 6 * N);
    d.tb.n_inv = c->ht.n_inv;
    d.tb.n_inv_s = c->ht.n_inv_s;
    d.tb.q = p->rlwe_q;
    d.tb.q2 = 2 * p->rlwe_q;
    d.tb.one_s = host::shoup_repr(1, p->rlwe_q);
    if (p->uitos_d) d.dec_uitos = make_decomp(p->rlwe_q, p->uitos_logb, p->uitos_d);
    d.dec_a = make_decomp(p->rlwe_q, p->rlrg_logb, p->rlrg_da);
    d.dec_b = make_decomp(p->rlwe_q, p->rlrg_logb, p->rlrg_db);
    d.dec_auto = make_decomp(p->rlwe_q, p->auto_logb, p->auto_d);
    d.dec_lwe = make_decomp(p->lwe_q, p->lwe_logb, p->lwe_d);
    d.da = p->rlrg_da;
    d.db = p->rlrg_db;
    d.dauto = p->auto_d;
    d.lwe_d = p->lwe_d;
    d.n_lwe = p->lwe_n;
    d.br_q = p->br_q;
    d.w = p->w;
    d.g = p->g;
    d.embedding = (2 * N) / p->br_q;
    d.lwe_q_log = host::ref_log2(p->lwe_q);
    d.lwe_mask = p->lwe_q - 1;
    d.lwe_qf = (double)p->lwe_q;
    d.rlwe_qf = (double)p->rlwe_q;
    d.qby4 = p->rlwe_q >> 2;
    d.br.q = p->rlwe_q;
    d.br.logq = 64 - (uint32_t)__builtin_clzll(p->rlwe_q + 1);
    d.br.alpha = d.br.logq + 3;
    host::u128 mu = ((host::u128)1 << (d.br.logq * 2 + 3)) / p->rlwe_q;
    d.br.mu_lo = (u64)mu;
    d.br.mu_hi = (u64)(mu >> 64);
    d.dlog = c->d_dlog;
    d.test_vec = c->d_tv;
    for (uint32_t k = 0; k <= p->w; k++) {
        int64_t el = k == 0 ? -(int64_t)p->g : 0;
        if (k > 0) {
            uint64_t gp = 1;
            for (uint32_t i = 0; i < k; i++) gp *= p->g;
            el = (int64_t)(gp % p->br_q);
        }
        d.auto_el[k] = host::auto_exp(el, N);
    }

    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    for (int var = BR_ONE; var < BR_NVAR; var++) {
        if (var != BR_ONE && !br_split(d)) continue;
        int per = 0;
        size_t lds = br_lds_bytes(logn, var != BR_ONE, p->br_q, p->lwe_n);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, blind_rotate_kernel(logn, var),
                                                         br_threads(logn, var), lds) != hipSuccess ||
            per < 1)
            per = 1;
        c->br_per_cu[var] = per;
    }
    for (int i = 0; i < 4; i++) hipEventCreate(&c->ev[i]);
    *out = c;
    return PZ_OK;
}

static void drop_graphs(pz_ctx *c)
{
    for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second);
    c->graphs.clear();
}

extern "C" int pz_ctx_destroy(pz_ctx *c)
{
    if (!c) return PZ_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graphs(c);
    pz_internal_exec_forget(c);
    hipFree(c->d_tables);
    hipFree(c->d_dlog);
    hipFree(c->d_tv);
    hipFree(c->d_digits);
    hipFree(c->d_ksb);
    hipFree(c->d_ks);
    hipFree(c->d_stats);
    hipFree(c->d_order);
    hipFree(c->ws);
    for (int i = 0; i < 4; i++)
        if (c->ev[i]) hipEventDestroy(c->ev[i]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return PZ_OK;
}

extern "C" int pz_ctx_info(pz_ctx *c, uint64_t *psi, uint32_t *num_cu, uint32_t *per_cu)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    if (psi) *psi = c->ht.psi;
    if (num_cu) *num_cu = (uint32_t)c->num_cu;
    if (per_cu) *per_cu = (uint32_t)c->br_per_cu[c->br_per_cu[BR_SPLIT] ? BR_SPLIT : BR_ONE];
    return PZ_OK;
}

extern "C" int pz_ctx_ntt_tables(pz_ctx *c, uint64_t *h)
{
    if (!c || !h) return fail(PZ_ERR_ARG, "null argument");
    uint32_t N = c->p.rlwe_n;
    memcpy(h, c->ht.psi_bo.data(), 8 * N);
    memcpy(h + N, c->ht.psi_bo_s.data(), 8 * N);
    memcpy(h + 2 * N, c->ht.ipsi_bo.data(), 8 * N);
    memcpy(h + 3 * N, c->ht.ipsi_bo_s.data(), 8 * N);
    h[4 * N] = c->ht.n_inv;
    h[4 * N + 1] = c->ht.n_inv_s;
    return PZ_OK;
}

extern "C" int pz_ctx_stream(pz_ctx *c, void **s)
{
    if (!c || !s) return fail(PZ_ERR_ARG, "null argument");
    *s = (void *)c->stream;
    return PZ_OK;
}

static hipStream_t S(pz_ctx *c, void *s) { return s ? (hipStream_t)s : c->stream; }

extern "C" int pz_malloc(pz_ctx *c, size_t bytes, void **p)
{
    if (!c || !p) return fail(PZ_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMalloc(p, bytes ? bytes : 8));
    return PZ_OK;
}

extern "C" int pz_free(pz_ctx *c, void *p)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    HIPCHK(hipFree(p));
    return PZ_OK;
}

extern "C" int pz_memcpy_htod(pz_ctx *c, void *d, const void *h, size_t bytes)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PZ_OK;
}

extern "C" int pz_memcpy_dtoh(pz_ctx *c, void *h, const void *d, size_t bytes)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    HIPCHK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PZ_OK;
}

extern "C" int pz_sync(pz_ctx *c)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipDeviceSynchronize());
    return PZ_OK;
}

static int ntt_common(pz_ctx *c, int mode, uint64_t *d, size_t nrows, void *s)
{
    if (!c || (!d && nrows)) return fail(PZ_ERR_ARG, "null argument");
    if (nrows == 0) return PZ_OK;
    launch_ntt(c->logn, mode, d, nrows, c->dc.tb, S(c, s));
    HIPCHK(hipGetLastError());
    return PZ_OK;
}

extern "C" int pz_ntt_forward_lazy(pz_ctx *c, uint64_t *d, size_t n, void *s) { return ntt_common(c, 0, d, n, s); }
extern "C" int pz_ntt_forward(pz_ctx *c, uint64_t *d, size_t n, void *s) { return ntt_common(c, 1, d, n, s); }
extern "C" int pz_ntt_backward_lazy(pz_ctx *c, uint64_t *d, size_t n, void *s) { return ntt_common(c, 2, d, n, s); }
extern "C" int pz_ntt_backward(pz_ctx *c, uint64_t *d, size_t n, void *s) { return ntt_common(c, 3, d, n, s); }

extern "C" int pz_shoup_matrix_fma(pz_ctx *c, uint64_t *out, const uint64_t *a,
                                   const uint64_t *as, const uint64_t *b, uint32_t d, size_t n,
                                   void *s)
{
    if (!c || !out || !a || !as || !b) return fail(PZ_ERR_ARG, "null argument");
    if (n == 0) return PZ_OK;
    launch_fma(out, a, as, b, d, n, c->p.rlwe_q, S(c, s));
    HIPCHK(hipGetLastError());
    return PZ_OK;
}

extern "C" int pz_decompose(pz_ctx *c, int which, const uint64_t *in, uint64_t *digits,
                            size_t count, void *s)
{
    if (!c || !in || !digits) return fail(PZ_ERR_ARG, "null argument");
    const Decomp *dc = nullptr;
    switch (which) {
    case 0: dc = &c->dc.dec_a; break;
    case 1: dc = &c->dc.dec_b; break;
    case 2: dc = &c->dc.dec_auto; break;
    case 3: dc = &c->dc.dec_lwe; break;
    default: return fail(PZ_ERR_ARG, "unknown decomposer");
    }
    if (count == 0) return PZ_OK;
    launch_decompose(in, digits, count, *dc, S(c, s));
    HIPCHK(hipGetLastError());
    return PZ_OK;
}

// Internal view of a context for the keygen translation unit.
int pz_internal_ctx_view(pz_ctx *c, pz_params *p, DevCtx *dc, int *logn, hipStream_t *s)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    if (hipSetDevice(c->device) != hipSuccess) return fail(PZ_ERR_HIP, "set device failed");
    *p = c->p;
    *dc = c->dc;
    *logn = c->logn;
    *s = c->stream;
    return PZ_OK;
}

// The context's workspace, at least `bytes` long (its contents are not kept
// across growth).
int pz_internal_workspace(pz_ctx *c, size_t bytes, void **out)
{
    if (bytes > c->ws_bytes) {
        if (c->stream) hipStreamSynchronize(c->stream);
        hipFree(c->ws);
        c->ws = nullptr;
        c->ws_bytes = 0;
        if (hipMalloc(&c->ws, bytes) != hipSuccess) return fail(PZ_ERR_HIP, "workspace allocation failed");
        c->ws_bytes = bytes;
    }
    *out = c->ws;
    return PZ_OK;
}

// Device bytes of a server key of this context's parameter set (what
// pz_key_bytes reports for a key uploaded to it).
size_t pz_internal_key_footprint(const pz_ctx *c)
{
    const pz_params &p = c->p;
    const size_t N = p.rlwe_n;
    const size_t rgsw_words = (size_t)p.lwe_n * (2 * (size_t)p.rlrg_da + 2 * (size_t)p.rlrg_db) * N;
    const size_t auto_words = (size_t)(p.w + 1) * 2 * (size_t)p.auto_d * N;
    const size_t kb = p.lwe_q <= 65536 ? 2 : 4, ldk = ((size_t)p.lwe_n + 1 + 63) / 64 * 64;
    return 16 * rgsw_words + 16 * auto_words + (kb + (PZ_KS_MFMA ? KS_PLANES : 0)) * N * p.lwe_d * ldk;
}

int pz_internal_ctx_device(const pz_ctx *c) { return c->device; }

// Build a key from device rows already in the reference layout (evaluation
// domain RGSW / auto rows, u64 KSK rows [N*d][n+1]): computes the Shoup twins
// and the packed KSK.  Takes ownership of rgsw and auto_k on success.
int pz_internal_key_adopt(pz_ctx *c, u64 *rgsw, u64 *auto_k, const u64 *ksk_rows, pz_key **out)
{
    const pz_params &p = c->p;
    uint32_t N = p.rlwe_n;
    size_t rgsw_rows = 2 * (size_t)p.rlrg_da + 2 * (size_t)p.rlrg_db;
    size_t auto_rows = 2 * (size_t)p.auto_d;
    size_t rgsw_words = (size_t)p.lwe_n * rgsw_rows * N;
    size_t auto_words = (size_t)(p.w + 1) * auto_rows * N;
    u32 kb = p.lwe_q <= 65536 ? 2 : 4;
    u32 n1 = p.lwe_n + 1, ldk = (n1 + 63) / 64 * 64;
    size_t ksk_dev = (size_t)N * p.lwe_d * ldk;
    pz_key *k = new pz_key();
    k->owner = c;
    bool ok = hipMalloc(&k->rgsw_s, 8 * rgsw_words) == hipSuccess &&
              hipMalloc(&k->auto_s, 8 * auto_words) == hipSuccess &&
              hipMalloc(&k->ksk, kb * ksk_dev) == hipSuccess &&
              (!PZ_KS_MFMA || hipMalloc(&k->ksk_planes, KS_PLANES * ksk_dev) == hipSuccess);
    if (!ok) {
        pz_key_free(k);
        return fail(PZ_ERR_HIP, "key allocation failed");
    }
    hipStream_t s = c->stream;
    // fold n^-1 into the RGSW and automorphism rows (DevKey)
    launch_scale(rgsw, rgsw_words, c->dc.tb.n_inv, c->dc.tb.n_inv_s, p.rlwe_q, s);
    launch_scale(auto_k, auto_words, c->dc.tb.n_inv, c->dc.tb.n_inv_s, p.rlwe_q, s);
    launch_shoup(rgsw, k->rgsw_s, rgsw_words, p.rlwe_q, s);
    launch_shoup(auto_k, k->auto_s, auto_words, p.rlwe_q, s);
    launch_pack_ksk(ksk_rows, k->ksk, (size_t)N * p.lwe_d, n1, ldk, kb, s);
    if (k->ksk_planes)
        launch_pack_ksk_planes(ksk_rows, k->ksk_planes, (size_t)N * p.lwe_d, n1, ldk, s);
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        pz_key_free(k);
        return fail(PZ_ERR_HIP, hipGetErrorString(e));
    }
    k->rgsw = rgsw;
    k->auto_k = auto_k;
    k->dk.rgsw = k->rgsw;
    k->dk.rgsw_s = k->rgsw_s;
    k->dk.auto_k = k->auto_k;
    k->dk.auto_s = k->auto_s;
    k->dk.ksk = k->ksk;
    k->dk.ksk_planes = k->ksk_planes;
    k->dk.rgsw_rows = (u32)rgsw_rows;
    k->dk.auto_rows = (u32)auto_rows;
    k->dk.ksk_bytes = kb;
    k->dk.ldk = ldk;
    k->bytes = 16 * rgsw_words + 16 * auto_words + kb * ksk_dev +
               (k->ksk_planes ? KS_PLANES * ksk_dev : 0);
    // pz_internal_key_footprint states the same sum
    *out = k;
    return PZ_OK;
}

extern "C" int pz_key_upload(pz_ctx *c, const uint64_t *h_rgsw, const uint64_t *h_auto,
                             const uint64_t *h_ksk, pz_key **out)
{
    if (!c || !h_rgsw || !h_auto || !h_ksk || !out) return fail(PZ_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const pz_params &p = c->p;
    uint32_t N = p.rlwe_n;
    size_t rgsw_words = (size_t)p.lwe_n * (2 * (size_t)p.rlrg_da + 2 * (size_t)p.rlrg_db) * N;
    size_t auto_words = (size_t)(p.w + 1) * 2 * (size_t)p.auto_d * N;
    size_t ksk_words = (size_t)N * p.lwe_d * (p.lwe_n + 1);
    u64 *rgsw = nullptr, *auto_k = nullptr, *tmp = nullptr;
    bool ok = hipMalloc(&rgsw, 8 * rgsw_words) == hipSuccess &&
              hipMalloc(&auto_k, 8 * auto_words) == hipSuccess &&
              hipMalloc(&tmp, 8 * ksk_words) == hipSuccess;
    if (ok) {
        hipStream_t s = c->stream;
        ok = hipMemcpyAsync(rgsw, h_rgsw, 8 * rgsw_words, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(auto_k, h_auto, 8 * auto_words, hipMemcpyHostToDevice, s) ==
                 hipSuccess &&
             hipMemcpyAsync(tmp, h_ksk, 8 * ksk_words, hipMemcpyHostToDevice, s) == hipSuccess;
    }
    int rc = ok ? pz_internal_key_adopt(c, rgsw, auto_k, tmp, out)
                : fail(PZ_ERR_HIP, "key upload failed");
    hipFree(tmp);
    if (rc != PZ_OK) {
        hipFree(rgsw);
        hipFree(auto_k);
    }
    return rc;
}

extern "C" int pz_key_free(pz_key *k)
{
    if (!k) return PZ_OK;
    if (k->owner) {
        hipStreamSynchronize(k->owner->stream);
        if (k->owner->graph_key_id == k->id) {
            drop_graphs(k->owner);
            k->owner->graph_key_id = 0;
        }
    }
    hipFree(k->rgsw);
    hipFree(k->rgsw_s);
    hipFree(k->auto_k);
    hipFree(k->auto_s);
    hipFree(k->ksk);
    hipFree(k->ksk_planes);
    delete k;
    return PZ_OK;
}

extern "C" int pz_key_download(pz_key *k, uint64_t *h_rgsw, uint64_t *h_auto, uint64_t *h_ksk)
{
    if (!k || !k->owner) return fail(PZ_ERR_ARG, "null key");
    pz_ctx *c = k->owner;
    const pz_params &p = c->p;
    uint32_t N = p.rlwe_n;
    size_t rgsw_words = (size_t)p.lwe_n * k->dk.rgsw_rows * N;
    size_t auto_words = (size_t)(p.w + 1) * k->dk.auto_rows * N;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    // device rows carry n^-1: multiply by N to return the key as generated
    auto unscale = [&](uint64_t *h, size_t words) {
        const uint64_t q = p.rlwe_q;
        for (size_t i = 0; i < words; i++)
            h[i] = (uint64_t)(((unsigned __int128)h[i] * N) % q);
    };
    if (h_rgsw) {
        HIPCHK(pz_copy_sync(h_rgsw, k->rgsw, 8 * rgsw_words, hipMemcpyDeviceToHost, c->stream));
        unscale(h_rgsw, rgsw_words);
    }
    if (h_auto) {
        HIPCHK(pz_copy_sync(h_auto, k->auto_k, 8 * auto_words, hipMemcpyDeviceToHost, c->stream));
        unscale(h_auto, auto_words);
    }
    if (h_ksk) {
        size_t rows = (size_t)N * p.lwe_d, n1 = p.lwe_n + 1, ld = k->dk.ldk;
        std::vector<uint8_t> packed(rows * ld * k->dk.ksk_bytes);
        HIPCHK(pz_copy_sync(packed.data(), k->ksk, packed.size(), hipMemcpyDeviceToHost,
                            c->stream));
        for (size_t r = 0; r < rows; r++)
            for (size_t col = 0; col < n1; col++) {
                size_t i = r * ld + col;
                h_ksk[r * n1 + col] = k->dk.ksk_bytes == 2
                                          ? reinterpret_cast<const uint16_t *>(packed.data())[i]
                                          : reinterpret_cast<const uint32_t *>(packed.data())[i];
            }
    }
    return PZ_OK;
}

extern "C" int pz_key_bytes(pz_key *k, size_t *bytes)
{
    if (!k || !bytes) return fail(PZ_ERR_ARG, "null argument");
    *bytes = k->bytes;
    return PZ_OK;
}

// Batched RLWE x RGSW (is_auto = false) or automorphism.  Where the split step
// runs (d = 1), the batch is ordered by key index on the device and the
// launch is persistent (one resident grid, each work group a contiguous run
// of the order), so consecutive RLWEs of a work group share key slices.
static int batched_rlwe_op(pz_ctx *c, pz_key *k, bool is_auto, uint64_t *rlwe,
                           const uint32_t *idx, const uint8_t *triv, size_t B, void *s)
{
    if (!c || !k || !rlwe || !idx || !triv) return fail(PZ_ERR_ARG, "null argument");
    KEY_OK(c, k);
    if (B == 0) return PZ_OK;
    hipStream_t st = S(c, s);
    if (!br_split(c->dc)) {
        launch_ext(c->logn, is_auto, rlwe, idx, triv, B, c->dc, k->dk, st);
        HIPCHK(hipGetLastError());
        return PZ_OK;
    }
    const u32 nkeys = is_auto ? c->p.w + 1 : c->p.lwe_n;
    const size_t need = B;
    if (need > c->order_cap) {
        // a launch on any stream may still read the old order
        HIPCHK(hipDeviceSynchronize());
        hipFree(c->d_order);
        c->d_order = nullptr;
        c->order_cap = 0;
        HIPCHK(hipMalloc(&c->d_order, 4 * need));
        c->order_cap = need;
    }
    u32 *order = launch_key_order(idx, B, nkeys, c->d_order, st) ? c->d_order : nullptr;
    const int grid = c->num_cu * step_split_per_cu(c->logn);
    launch_ext(c->logn, is_auto, rlwe, idx, triv, B, c->dc, k->dk, st, order, grid);
    HIPCHK(hipGetLastError());
    return PZ_OK;
}

extern "C" int pz_rlwe_by_rgsw(pz_ctx *c, pz_key *k, uint64_t *rlwe, const uint32_t *idx,
                               const uint8_t *triv, size_t B, void *s)
{
    return batched_rlwe_op(c, k, false, rlwe, idx, triv, B, s);
}

extern "C" int pz_rlwe_auto(pz_ctx *c, pz_key *k, uint64_t *rlwe, const uint32_t *idx,
                            const uint8_t *triv, size_t B, void *s)
{
    return batched_rlwe_op(c, k, true, rlwe, idx, triv, B, s);
}

// grow the PBS pipeline scratch to hold B ciphertexts
static int ensure_scratch(pz_ctx *c, size_t B)
{
    if (B <= c->cap) return PZ_OK;
    drop_graphs(c);  // captured graphs reference the old buffers
    hipStreamSynchronize(c->stream);
    hipFree(c->d_digits);
    hipFree(c->d_ksb);
    hipFree(c->d_ks);
    hipFree(c->d_stats);
    c->d_stats = nullptr;
    c->d_digits = nullptr;
    c->d_ksb = nullptr;
    c->d_ks = nullptr;
    size_t rows = (size_t)c->p.rlwe_n * c->p.lwe_d;
    size_t cap = B < 64 ? 64 : B;
    cap = (cap + KS_CT - 1) / KS_CT * KS_CT;  // the key switch reads whole ciphertext tiles
    HIPCHK(hipMalloc(&c->d_digits, rows * cap));
    HIPCHK(hipMalloc(&c->d_ksb, 4 * cap));
    HIPCHK(hipMalloc(&c->d_ks, 4 * cap * (c->p.lwe_n + 1)));
    HIPCHK(hipMalloc(&c->d_stats, 4 * cap));
    c->cap = cap;
    return PZ_OK;
}

// Executor view: grows the pipeline scratch to B ciphertexts and exposes it.
int pz_internal_exec_view(pz_ctx *c, size_t B, pz_exec_view *v, DevCtx *dc)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    int rc = ensure_scratch(c, B);
    if (rc) return rc;
    v->p = c->p;
    v->logn = c->logn;
    v->digits = c->d_digits;
    v->ksb = c->d_ksb;
    v->ks = c->d_ks;
    v->stats = c->d_stats;
    size_t resident = (size_t)c->num_cu * c->br_per_cu[BR_ONE];
    v->grid = (int)(resident < 1 ? 1 : resident);
    *dc = c->dc;
    return PZ_OK;
}

const DevKey *pz_internal_devkey(const pz_key *k) { return &k->dk; }

// Contexts whose keys are interchangeable: same device, same parameter set.
bool pz_internal_ctx_compatible(const pz_ctx *a, const pz_ctx *b)
{
    if (!a || !b) return false;
    if (a == b) return true;
    const pz_params &x = a->p, &y = b->p;
    return a->device == b->device && x.rlwe_q == y.rlwe_q && x.rlwe_n == y.rlwe_n &&
           x.lwe_n == y.lwe_n && x.lwe_q == y.lwe_q && x.br_q == y.br_q &&
           x.lwe_logb == y.lwe_logb && x.lwe_d == y.lwe_d && x.rlrg_logb == y.rlrg_logb &&
           x.rlrg_da == y.rlrg_da && x.rlrg_db == y.rlrg_db && x.auto_logb == y.auto_logb &&
           x.auto_d == y.auto_d && x.g == y.g && x.w == y.w &&
           x.lwe_secret_gaussian == y.lwe_secret_gaussian && x.uitos_logb == y.uitos_logb &&
           x.uitos_d == y.uitos_d && x.rgrg_logb == y.rgrg_logb && x.rgrg_da == y.rgrg_da &&
           x.rgrg_db == y.rgrg_db;
}

// A key serves any context of its parameter set on its device.
bool pz_internal_key_fits(const pz_ctx *c, const pz_key *k)
{
    return k && pz_internal_ctx_compatible(c, k->owner);
}


extern "C" int pz_lwe_key_switch(pz_ctx *c, pz_key *k, const uint64_t *in, uint64_t *out,
                                 size_t B, void *s)
{
    if (!c || !k || !in || !out) return fail(PZ_ERR_ARG, "null argument");
    KEY_OK(c, k);
    if (B == 0) return PZ_OK;
    int rc = ensure_scratch(c, B);
    if (rc) return rc;
    hipStream_t st = S(c, s);
    const pz_params &p = c->p;
    u32 n1 = p.lwe_n + 1;
    const bool rm = ks_rowmajor(k->dk);  // the MFMA key switch's digit layout
    launch_prep(PREP_KSONLY, in, nullptr, nullptr, nullptr, nullptr, p.rlwe_n + 1, B,
                c->d_digits, c->d_ksb, c->d_ks, c->dc, st, rm);
    launch_ks(c->d_digits, c->d_ksb, k->dk.ksk, k->dk.ksk_bytes, p.rlwe_n * p.lwe_d, n1,
              k->dk.ldk, 0, B, c->d_ks, st, rm ? k->dk.ksk_planes : nullptr);
    launch_widen(c->d_ks, out, B * n1, (u32)(p.lwe_q - 1), st);
    HIPCHK(hipGetLastError());
    return PZ_OK;
}

// Blind-rotation kernel variant and grid for a batch of B bootstraps.  With
// d = 1 the split kernel (two half work groups per bootstrap) is used: a
// latency variant while every bootstrap has a CU of its own (4 coefficients
// per thread, or 8 in mode 2), a throughput variant with two bootstraps per
// CU beyond that (8 coefficients per thread, or 4 in mode 3).  Otherwise, or
// in mode 1, one work group per bootstrap.  An autotuned shape for the
// batch's bucket overrides this in mode 0.
static int batch_bucket(size_t B)
{
    int b = 0;
    while (((size_t)1 << b) < B && b < 39) b++;
    return b;
}

static void br_plan(const pz_ctx *c, size_t B, int *var, int *grid)
{
    const pz_ctx::Tuned &t = c->tuned[batch_bucket(B)];
    if (c->br_mode == 0 && t.var >= 0) {
        const size_t res = (size_t)c->num_cu * t.per_cu;
        *var = t.var;
        *grid = (int)(B < res ? (B > 0 ? B : 1) : res);
        return;
    }
    int v = BR_ONE;
    if (c->br_per_cu[BR_SPLIT] > 0 && c->br_mode != 1) {
        // automatic: the 4-coefficient latency variant (twice the waves of a
        // bootstrap, measured faster for one bootstrap per CU), the
        // 8-coefficient throughput variant (two bootstraps per CU)
        const int lat = c->br_mode == 2 ? BR_SPLIT_LAT : BR_SPLIT4_LAT;
        const int thr = c->br_mode == 3 ? BR_SPLIT4 : BR_SPLIT;
        v = B <= (size_t)c->num_cu * c->br_per_cu[lat] ? lat : thr;
    }
    const size_t res = (size_t)c->num_cu * (c->br_per_cu[v] > 0 ? c->br_per_cu[v] : 1);
    *var = v;
    *grid = (int)(B < res ? (B > 0 ? B : 1) : res);
}

int pz_internal_br_plan(pz_ctx *c, size_t B, int *var, int *grid)
{
    br_plan(c, B, var, grid);
    return PZ_OK;
}

extern "C" int pz_set_blind_rotate_mode(pz_ctx *c, int mode)
{
    if (!c || mode < 0 || mode > 3) return fail(PZ_ERR_ARG, "mode must be 0, 1, 2 or 3");
    if (mode >= 2 && c->br_per_cu[BR_SPLIT] == 0)
        return fail(PZ_ERR_UNSUPPORTED, "the split kernel needs d_a = d_b = d_auto = 1");
    if (mode != c->br_mode) {
        HIPCHK(hipSetDevice(c->device));
        hipStreamSynchronize(c->stream);
        drop_graphs(c);  // captured graphs hold the previous kernel choice
        c->br_mode = mode;
    }
    return PZ_OK;
}

// moddown/KS/blind-rotate sequence on stream st; var >= 0 forces the
// blind-rotation variant and grid (the autotuner's candidates)
static void pbs_sequence(pz_ctx *c, pz_key *k, int mode, const uint64_t *c0, const uint64_t *c1,
                         const uint32_t *ops, uint64_t *out, size_t B, hipStream_t st,
                         int force_var = -1, int force_grid = 0, const PbsScratch *sc = nullptr,
                         bool events = true)
{
    const pz_params &p = c->p;
    u32 n1 = p.lwe_n + 1;
    const PbsScratch own{c->d_digits, c->d_ksb, c->d_ks, c->d_stats};
    if (!sc) sc = &own;  // the context's grow-on-demand scratch
    events = events && c->timing;
    const bool rm = ks_rowmajor(k->dk);  // the MFMA key switch's digit layout
    if (events) hipEventRecord(c->ev[0], st);
    launch_prep(mode, c0, c1, nullptr, nullptr, ops, p.rlwe_n + 1, B, sc->digits, sc->ksb,
                sc->ks, c->dc, st, rm);
    if (events) hipEventRecord(c->ev[1], st);
    launch_ks(sc->digits, sc->ksb, k->dk.ksk, k->dk.ksk_bytes, p.rlwe_n * p.lwe_d, n1,
              k->dk.ldk, 0, B, sc->ks, st, rm ? k->dk.ksk_planes : nullptr);
    if (events) hipEventRecord(c->ev[2], st);
    int var = BR_ONE, grid = 1;
    if (force_var >= 0) {
        var = force_var;
        grid = force_grid;
    } else {
        br_plan(c, B, &var, &grid);
    }
    launch_blind_rotate(c->logn, var, grid, sc->ks, ops, B, out, sc->stats, c->dc, k->dk,
                        nullptr, nullptr, nullptr, st);
    if (events) hipEventRecord(c->ev[3], st);
}

// Timing events as nodes of a captured gate graph: the capture holds the
// three kernels (prep, key switch, blind rotation) in one chain; an event
// record node goes in front of the first and after each, and each kernel
// also waits for the record in front of it.  (Events recorded on a
// capturing stream cannot be timed.)
static hipError_t add_timing_nodes(pz_ctx *c, hipGraph_t g)
{
    size_t nn = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &nn);
    if (e != hipSuccess) return e;
    if (nn != 3) return hipErrorNotSupported;
    hipGraphNode_t nodes[3], chain[3];
    e = hipGraphGetNodes(g, nodes, &nn);
    if (e != hipSuccess) return e;
    // order the chain: the node that depends on the previous one comes next
    hipGraphNode_t prev = nullptr;
    for (int pos = 0; pos < 3; pos++) {
        chain[pos] = nullptr;
        for (int i = 0; i < 3; i++) {
            size_t nd = 0;
            hipGraphNode_t dep = nullptr;
            if ((e = hipGraphNodeGetDependencies(nodes[i], nullptr, &nd)) != hipSuccess) return e;
            if (nd > 1) return hipErrorNotSupported;
            if (nd == 1 && (e = hipGraphNodeGetDependencies(nodes[i], &dep, &nd)) != hipSuccess)
                return e;
            if ((prev == nullptr && nd == 0) || (prev != nullptr && nd == 1 && dep == prev))
                chain[pos] = nodes[i];
        }
        if (!chain[pos]) return hipErrorNotSupported;
        hipGraphNodeType t;
        if ((e = hipGraphNodeGetType(chain[pos], &t)) != hipSuccess) return e;
        if (t != hipGraphNodeTypeKernel) return hipErrorNotSupported;
        prev = chain[pos];
    }
    hipGraphNode_t rec[4];
    for (int i = 0; i < 4; i++) {
        const hipGraphNode_t *dep = i ? &chain[i - 1] : nullptr;
        if ((e = hipGraphAddEventRecordNode(&rec[i], g, dep, i ? 1 : 0, c->ev[i])) != hipSuccess)
            return e;
        if (i < 3 && (e = hipGraphAddDependencies(g, &rec[i], &chain[i], 1)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

static int collect_timing(pz_ctx *c, hipStream_t st)
{
    if (!c->timing) return PZ_OK;
    HIPCHK(hipEventSynchronize(c->ev[3]));
    HIPCHK(hipEventElapsedTime(&c->t_prep, c->ev[0], c->ev[1]));
    HIPCHK(hipEventElapsedTime(&c->t_ks, c->ev[1], c->ev[2]));
    HIPCHK(hipEventElapsedTime(&c->t_br, c->ev[2], c->ev[3]));
    (void)st;
    return PZ_OK;
}

// Autotuner: times every blind-rotation shape that fits the device (kernel
// variant x resident work groups per CU) on a batch of synthetic gates with
// this key and keeps the fastest for B's bucket.  Batches larger than four
// rounds of the largest resident grid are sampled at that size (every shape
// is persistent there).  Inputs are uniform words mod Q (any word is a valid
// PBS input); outputs are discarded.
extern "C" int pz_autotune(pz_ctx *c, pz_key *k, size_t B_req, uint32_t *h_choice)
{
    if (!c || !k || B_req == 0) return fail(PZ_ERR_ARG, "null argument or empty batch");
    KEY_OK(c, k);
    HIPCHK(hipSetDevice(c->device));
    int max_per = 1;
    for (int v = 0; v < BR_NVAR; v++) max_per = std::max(max_per, c->br_per_cu[v]);
    const size_t sample = 4 * (size_t)c->num_cu * max_per;
    const size_t B = B_req < sample ? B_req : sample;
    int rc = ensure_scratch(c, B);
    if (rc) return rc;
    const size_t N1 = (size_t)c->p.rlwe_n + 1;
    std::vector<u64> h_in(2 * B * N1);
    std::vector<u32> h_ops(B);
    u64 x = 0x9e3779b97f4a7c15ull;
    for (auto &v : h_in) {  // splitmix64 words reduced mod Q
        x += 0x9e3779b97f4a7c15ull;
        u64 z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        v = (z ^ (z >> 31)) % c->p.rlwe_q;
    }
    for (size_t i = 0; i < B; i++) h_ops[i] = (i & 1) ? 4u : 0u;  // NAND / XOR
    u64 *d_in = nullptr, *d_out = nullptr;
    u32 *d_ops = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipMalloc(&d_in, 8 * h_in.size()) == hipSuccess &&
              hipMalloc(&d_out, 8 * B * N1) == hipSuccess &&
              hipMalloc(&d_ops, 4 * B) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess &&
              hipMemcpy(d_in, h_in.data(), 8 * h_in.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_ops, h_ops.data(), 4 * B, hipMemcpyHostToDevice) == hipSuccess;
    int best_var = -1, best_per = 0;
    float best_ms = 0;
    for (int v = 0; ok && v < BR_NVAR; v++) {
        if (c->br_per_cu[v] <= 0) continue;
        for (int per = 1; ok && per <= c->br_per_cu[v]; per++) {
            const size_t res = (size_t)c->num_cu * per;
            const int grid = (int)(B < res ? B : res);
            float ms = 0;
            for (int rep = 0; rep < 3 && ok; rep++) {  // one warm-up, two timed
                ok = hipEventRecord(e0, c->stream) == hipSuccess;
                pbs_sequence(c, k, PREP_GATE, d_in, d_in + B * N1, d_ops, d_out, B, c->stream, v,
                             grid);
                ok = ok && hipGetLastError() == hipSuccess &&
                     hipEventRecord(e1, c->stream) == hipSuccess &&
                     hipEventSynchronize(e1) == hipSuccess;
                float t = 0;
                if (ok && rep > 0 && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms += t;
            }
            if (ok && (best_var < 0 || ms < best_ms)) {
                best_var = v;
                best_per = per;
                best_ms = ms;
            }
        }
    }
    hipFree(d_in);
    hipFree(d_out);
    hipFree(d_ops);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (!ok || best_var < 0) return fail(PZ_ERR_HIP, "autotune run failed");
    pz_ctx::Tuned &t = c->tuned[batch_bucket(B_req)];
    t.var = best_var;
    t.per_cu = best_per;
    drop_graphs(c);  // captured graphs hold the previous choice
    if (h_choice) {
        h_choice[0] = (uint32_t)best_var;
        h_choice[1] = (uint32_t)best_per;
        h_choice[2] = (uint32_t)(best_ms / 2 * 1000);  // microseconds per batch
    }
    return PZ_OK;
}

extern "C" int pz_pbs(pz_ctx *c, pz_key *k, const uint32_t *gate, uint64_t *lwe, size_t B,
                      void *s)
{
    if (!c || !k || !gate || !lwe) return fail(PZ_ERR_ARG, "null argument");
    KEY_OK(c, k);
    if (B == 0) return PZ_OK;
    int rc = ensure_scratch(c, B);
    if (rc) return rc;
    hipStream_t st = S(c, s);
    pbs_sequence(c, k, PREP_PBS, lwe, nullptr, gate, lwe, B, st);
    HIPCHK(hipGetLastError());
    return collect_timing(c, st);
}

extern "C" int pz_gates(pz_ctx *c, pz_key *k, const uint32_t *ops, const uint64_t *c0,
                        const uint64_t *c1, uint64_t *out, size_t B, void *s)
{
    if (!c || !k || !ops || !c0 || !c1 || !out) return fail(PZ_ERR_ARG, "null argument");
    KEY_OK(c, k);
    if (B == 0) return PZ_OK;
    int rc = ensure_scratch(c, B);
    if (rc) return rc;
    hipStream_t st = S(c, s);
    if (c->graph_key_id != k->id) {
        drop_graphs(c);
        c->graph_key_id = k->id;
    }
    // with timing on, the graph carries event-record nodes around each kernel,
    // so the per-kernel times come from the same replays as the batch's
    GraphKey gk{ops, c0, c1, out, B, c->timing};
    auto it = c->graphs.find(gk);
    if (it == c->graphs.end()) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        pbs_sequence(c, k, PREP_GATE, c0, c1, ops, out, B, c->stream, -1, 0, nullptr, false);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess) return fail(PZ_ERR_HIP, hipGetErrorString(e));
        if (c->timing && (e = add_timing_nodes(c, g)) != hipSuccess) {
            hipGraphDestroy(g);
            return fail(PZ_ERR_HIP, (std::string("timing nodes: ") + hipGetErrorString(e)).c_str());
        }
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(PZ_ERR_HIP, hipGetErrorString(e));
        if (c->graphs.size() > 32) drop_graphs(c);
        it = c->graphs.emplace(gk, ge).first;
    }
    HIPCHK(hipGraphLaunch(it->second, st));
    return collect_timing(c, st);
}

// pz_gates without any allocation: the pipeline scratch is this device's
// fixed array (no grow-on-demand buffers, no captured graph), so a call holds
// at most noalloc_max_batch bootstraps.  Every call on a device uses the same
// array; each kernel writes the parts it reads before reading them (the prep
// kernel clears the key-switch outputs), so no call depends on what an
// earlier one left, but calls that overlap on one device race.
extern "C" int pz_gates_noalloc(pz_ctx *c, pz_key *k, const uint32_t *ops, const uint64_t *c0,
                                const uint64_t *c1, uint64_t *out, size_t B, void *s)
{
    if (!c || !k || !ops || !c0 || !c1 || !out) return fail(PZ_ERR_ARG, "null argument");
    KEY_OK(c, k);
    if (B == 0) return PZ_OK;
    const size_t bmax = noalloc_max_batch(c->p);
    if (B > bmax) {
        char msg[160];
        snprintf(msg, sizeof(msg),
                 "batch of %zu exceeds the fixed scratch: at most %zu bootstraps per call for "
                 "this parameter set (%zu MiB per device)",
                 B, bmax, NOALLOC_SCRATCH_BYTES >> 20);
        return fail(PZ_ERR_ARG, msg);
    }
    if (c->device >= NOALLOC_MAX_DEVICES || !g_scratch_ptr[c->device])
        return fail(PZ_ERR_UNSUPPORTED, "no fixed scratch on this device");
    PbsScratch sc;
    const size_t cap = (B + KS_CT - 1) / KS_CT * KS_CT;
    noalloc_layout(c->p, cap, (uint8_t *)g_scratch_ptr[c->device], &sc);
    hipStream_t st = S(c, s);
    pbs_sequence(c, k, PREP_GATE, c0, c1, ops, out, B, st, -1, 0, &sc);
    HIPCHK(hipGetLastError());
    return collect_timing(c, st);
}

extern "C" int pz_not(pz_ctx *c, uint64_t *d, size_t B, void *s)
{
    if (!c || !d) return fail(PZ_ERR_ARG, "null argument");
    if (B == 0) return PZ_OK;
    launch_not(d, B * (c->p.rlwe_n + 1), c->p.rlwe_q, S(c, s));
    HIPCHK(hipGetLastError());
    return PZ_OK;
}

extern "C" int pz_last_auto_counts(pz_ctx *c, uint32_t *h_counts, size_t B)
{
    if (!c || !h_counts) return fail(PZ_ERR_ARG, "null argument");
    if (B > c->cap) return fail(PZ_ERR_ARG, "more counts than the last batch holds");
    HIPCHK(hipMemcpyAsync(h_counts, c->d_stats, 4 * B, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PZ_OK;
}

extern "C" int pz_set_timing(pz_ctx *c, int enabled)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    c->timing = enabled != 0;
    return PZ_OK;
}

extern "C" int pz_last_timing(pz_ctx *c, float *a, float *b, float *d)
{
    if (!c) return fail(PZ_ERR_ARG, "null context");
    if (a) *a = c->t_prep;
    if (b) *b = c->t_ks;
    if (d) *d = c->t_br;
    return PZ_OK;
}

// VALU issue-rate probe (no counterpart in the reference; measures the chip's
// integer VALU rates for the roofline of the PBS kernels).
extern "C" int pz_valu_probe(pz_ctx *c, int kind, uint32_t iters, double *ops_per_s)
{
    if (!c || !ops_per_s || kind < 0 || kind > 5 || iters == 0)
        return fail(PZ_ERR_ARG, "bad probe arguments");
    HIPCHK(hipSetDevice(c->device));
    const int grid = c->num_cu * 16;  // 16 work groups of 4 waves per CU: 2 rounds at 8 waves/SIMD
    const size_t threads = (size_t)grid * 256;
    u64 *out = nullptr;
    HIPCHK(hipMalloc(&out, 8 * threads));
    const u64 q = c->p.rlwe_q, w = (q >> 1) | 1;
    const u64 ws = (u64)(((unsigned __int128)w << 64) / q);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    launch_valu_probe(kind, grid, out, iters / 8 + 1, 0x9e3779b1u, w, ws, q, c->stream);  // warm-up
    hipEventRecord(e0, c->stream);
    launch_valu_probe(kind, grid, out, iters, 0x9e3779b1u, w, ws, q, c->stream);
    hipEventRecord(e1, c->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(out);
    if (e != hipSuccess) return fail(PZ_ERR_HIP, hipGetErrorString(e));
    *ops_per_s = (double)threads * iters * PROBE_CHAINS / (ms * 1e-3);
    return PZ_OK;
}
