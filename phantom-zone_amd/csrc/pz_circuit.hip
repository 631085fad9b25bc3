// pz_circuit.hip -- gate-DAG circuits (FheUint8 arithmetic) evaluated level by
// level: every ready gate of a level, from any number of independent
// operations, is bootstrapped in one batched launch.
//
// The circuit builders restate the reference's FheUint8 circuits over the
// BooleanGates seam (src/shortint/ops.rs:5-356, operators in
// src/shortint/mod.rs:30-293) with the same gate order and operand order, so
// every output ciphertext equals the reference's serial evaluation.
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <queue>
#include <string>
#include <vector>

#include "../../include/pz.h"
#include "pz_exec.hpp"

namespace pz {

// ---------------------------------------------------------------------------
// Executor
// ---------------------------------------------------------------------------
__global__ void k_not_rows(u64 *pool, const u32 *in, const u32 *out, size_t count, u32 N1, u64 q)
{
    size_t total = count * N1;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        size_t g = t / N1;
        u32 col = (u32)(t % N1);
        pool[(size_t)out[g] * N1 + col] = neg_mod(pool[(size_t)in[g] * N1 + col], q);
    }
}

__global__ void k_copy_rows(u64 *pool, const u64 *src, const u32 *dst_rows, size_t count, u32 N1)
{
    size_t total = count * N1;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        size_t g = t / N1;
        u32 col = (u32)(t % N1);
        pool[(size_t)dst_rows[g] * N1 + col] = src[g * N1 + col];
    }
}

__global__ void k_gather_rows(const u64 *pool, u64 *dst, const u32 *src_rows, size_t count, u32 N1)
{
    size_t total = count * N1;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        size_t g = t / N1;
        u32 col = (u32)(t % N1);
        dst[g * N1 + col] = pool[(size_t)src_rows[g] * N1 + col];
    }
}

static int grid_rows(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 65535 * 8); }

// Pinned host staging + device copies of the per-batch index arrays.
struct Staging {
    u32 *h = nullptr, *d = nullptr;
    DevKey *hk = nullptr, *dk = nullptr;
    size_t cap = 0, kcap = 0;
    ~Staging()
    {
        if (h) (void)hipHostFree(h);
        if (d) (void)hipFree(d);
        if (hk) (void)hipHostFree(hk);
        if (dk) (void)hipFree(dk);
    }
    bool reserve(size_t words, size_t keys)
    {
        if (words > cap) {
            if (h) (void)hipHostFree(h);
            if (d) (void)hipFree(d);
            h = nullptr;
            d = nullptr;
            size_t c = std::max<size_t>(words, 4096);
            if (hipHostMalloc(&h, 4 * c) != hipSuccess || hipMalloc(&d, 4 * c) != hipSuccess)
                return false;
            cap = c;
        }
        if (keys > kcap) {
            if (hk) (void)hipHostFree(hk);
            if (dk) (void)hipFree(dk);
            hk = nullptr;
            dk = nullptr;
            size_t c = std::max<size_t>(keys, 64);
            if (hipHostMalloc(&hk, sizeof(DevKey) * c) != hipSuccess ||
                hipMalloc(&dk, sizeof(DevKey) * c) != hipSuccess)
                return false;
            kcap = c;
        }
        return true;
    }
};

// Two staging areas per context, used in turn: a batch's index arrays stay
// untouched until the batch has finished (an event recorded after its last
// launch), so the host prepares and queues the next batch while the previous
// one runs.  Batches of different contexts (a server serving several
// parameter sets) run on their own streams with their own staging.
struct StagingPair {
    Staging slot[2];
    hipEvent_t done[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int next = 0;
    ~StagingPair()
    {
        for (int i = 0; i < 2; i++)
            if (done[i]) {
                (void)hipEventSynchronize(done[i]);
                (void)hipEventDestroy(done[i]);
            }
    }
};
static std::mutex g_stage_mu;
static std::map<const pz_ctx *, StagingPair *> g_stages;

static StagingPair &stage_of(const pz_ctx *c)
{
    std::lock_guard<std::mutex> lk(g_stage_mu);
    StagingPair *&st = g_stages[c];
    if (!st) st = new StagingPair();
    return *st;
}

}  // namespace pz

void pz_internal_exec_forget(const pz_ctx *c)
{
    using pz::g_stage_mu;
    using pz::g_stages;
    pz::StagingPair *st = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        auto it = g_stages.find(c);
        if (it == g_stages.end()) return;
        st = it->second;
        g_stages.erase(it);
    }
    delete st;  // the context's device is current (pz_ctx_destroy)
}

namespace pz {

int exec_batch(pz_ctx *c, const std::vector<const pz_key *> &keys, std::vector<GateRec> &recs,
               u64 *pool, hipStream_t st)
{
    std::vector<GateRec> pbs, nots;
    for (auto &r : recs) (r.op <= 5 ? pbs : nots).push_back(r);
    std::stable_sort(pbs.begin(), pbs.end(),
                     [](const GateRec &a, const GateRec &b) { return a.key < b.key; });
    const size_t B = pbs.size();
    pz_exec_view v;
    DevCtx dc;
    int rc = pz_internal_exec_view(c, B ? B : 1, &v, &dc);
    if (rc) return rc;
    StagingPair &pair = stage_of(c);
    const int slot = pair.next;
    Staging &g_stage = pair.slot[slot];
    const u32 N1 = v.p.rlwe_n + 1, n1 = v.p.lwe_n + 1;
    // staging: ops | in0 | in1 | out | key | not_in | not_out | KS work list
    const size_t nn = nots.size();
    // key-switch work list for batches that mix keys: one entry per
    // (ciphertext tile, key segment) -- a tile that straddles two keys'
    // segments appears once for each
    std::vector<u32> work;
    if (keys.size() > 1) {
        size_t s0 = 0;
        while (s0 < B) {
            size_t e0 = s0;
            while (e0 < B && pbs[e0].key == pbs[s0].key) e0++;
            for (size_t t = s0 / KS_CT; t <= (e0 - 1) / KS_CT; t++) {
                work.push_back((u32)t);
                work.push_back((u32)s0);
                work.push_back((u32)e0);
                work.push_back(pbs[s0].key);
            }
            s0 = e0;
        }
    }
    const size_t nw = work.size();
    // the batch before last used this slot (its copies read the pinned
    // buffer, its kernels the device copy): wait for it to finish
    if (!pair.done[slot] && hipEventCreateWithFlags(&pair.done[slot], hipEventDisableTiming) !=
                                hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipEventCreate failed");
    if (pair.used[slot] && hipEventSynchronize(pair.done[slot]) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipEventSynchronize failed");
    if (!g_stage.reserve(5 * B + 2 * nn + nw + 8, keys.size()))
        return PZ_FAIL(PZ_ERR_HIP, "g_stage.reserve failed");
    u32 *h = g_stage.h;
    for (size_t i = 0; i < B; i++) {
        h[i] = pbs[i].op;
        h[B + i] = pbs[i].in0;
        h[2 * B + i] = pbs[i].in1;
        h[3 * B + i] = pbs[i].out;
        h[4 * B + i] = pbs[i].key;
    }
    for (size_t i = 0; i < nn; i++) {
        h[5 * B + i] = nots[i].in0;
        h[5 * B + nn + i] = nots[i].out;
    }
    if (nw) memcpy(h + 5 * B + 2 * nn, work.data(), 4 * nw);
    for (size_t k = 0; k < keys.size(); k++) g_stage.hk[k] = *pz_internal_devkey(keys[k]);
    u32 *d = g_stage.d;
    if (hipMemcpyAsync(d, h, 4 * (5 * B + 2 * nn + nw), hipMemcpyHostToDevice, st) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipMemcpyAsync failed");
    if (keys.size() > 1 &&
        hipMemcpyAsync(g_stage.dk, g_stage.hk, sizeof(DevKey) * keys.size(),
                       hipMemcpyHostToDevice, st) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipMemcpyAsync failed");
    if (B) {
        const DevKey *k0 = pz_internal_devkey(keys[pbs[0].key]);
        // one key: the MFMA key switch when the key has its planes; mixed keys:
        // the per-tile VALU kernel (tile-major digits)
        const bool rm = keys.size() == 1 && ks_rowmajor(*k0);
        launch_prep(PREP_GATE, pool, pool, d + B, d + 2 * B, d, N1, B, v.digits, v.ksb, v.ks, dc,
                    st, rm);
        if (keys.size() > 1)
            launch_ks_multi(v.digits, v.ksb, g_stage.dk, d + 5 * B + 2 * nn, nw / 4,
                            k0->ksk_bytes, v.p.rlwe_n * v.p.lwe_d, n1, k0->ldk, v.ks, st);
        else
            launch_ks(v.digits, v.ksb, k0->ksk, k0->ksk_bytes, v.p.rlwe_n * v.p.lwe_d, n1,
                      k0->ldk, 0, B, v.ks, st, rm ? k0->ksk_planes : nullptr);
        int var = BR_ONE, grid = 1;
        pz_internal_br_plan(c, B, &var, &grid);
        launch_blind_rotate(v.logn, var, grid, v.ks, d, B, pool, v.stats, dc, *k0,
                            keys.size() > 1 ? g_stage.dk : nullptr,
                            keys.size() > 1 ? d + 4 * B : nullptr, d + 3 * B, st);
    }
    // NOT gates in waves: a NOT whose input is written by an earlier NOT of
    // this batch goes to a later wave
    if (nn) {
        std::vector<int> wave(nn, 0);
        int maxw = 0;
        for (size_t i = 0; i < nn; i++)
            for (size_t j = 0; j < i; j++)
                if (nots[j].out == nots[i].in0) {
                    wave[i] = std::max(wave[i], wave[j] + 1);
                    maxw = std::max(maxw, wave[i]);
                }
        if (maxw == 0) {
            hipLaunchKernelGGL(k_not_rows, grid_rows(nn * N1), 256, 0, st, pool, d + 5 * B,
                               d + 5 * B + nn, nn, N1, v.p.rlwe_q);
        } else {
            for (int w = 0; w <= maxw; w++) {
                for (size_t i = 0; i < nn; i++) {
                    if (wave[i] != w) continue;
                    hipLaunchKernelGGL(k_not_rows, 1, 256, 0, st, pool, d + 5 * B + i,
                                       d + 5 * B + nn + i, (size_t)1, N1, v.p.rlwe_q);
                }
            }
        }
    }
    rc = PZ_LAUNCH_STATUS();
    if (rc) return rc;
    if (hipEventRecord(pair.done[slot], st) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipEventRecord failed");
    pair.used[slot] = true;
    pair.next = slot ^ 1;
    return PZ_OK;
}

// ---------------------------------------------------------------------------
// Circuit (gate DAG)
// ---------------------------------------------------------------------------
struct Node {
    u32 kind;  // 0 input, 1 bootstrapped gate, 2 NOT
    u32 op, a, b;
    u32 level;
};

}  // namespace pz

using namespace pz;

struct pz_circuit {
    pz_ctx *ctx = nullptr;
    std::vector<Node> nodes;
    std::vector<u32> inputs;  // input wires in order
    int div_flag = -1;        // running div-by-zero flag wire
    u64 *pool = nullptr;
    size_t pool_rows = 0;
    size_t ran_nodes = 0;     // wires the last run computed (0: never run)
    u32 N1 = 0;
    u32 depth = 0, n_pbs = 0;
    std::vector<std::vector<GateRec>> levels;  // schedule (built lazily)
    bool scheduled = false;
    // launch plan of the schedule, one device copy: per level the bootstrapped
    // gates' ops | in0 | in1 | out, then each NOT wave's in | out
    struct Wave {
        size_t off;
        u32 count;
    };
    struct LevelPlan {
        size_t off;
        u32 B;
        std::vector<Wave> waves;
    };
    std::vector<u32> plan;
    std::vector<LevelPlan> lplan;
    u32 max_B = 0;
    u32 *d_plan = nullptr;
    size_t d_plan_cap = 0;
    bool plan_on_device = false;
};

typedef u32 W;

struct Builder {
    pz_circuit *c;
    W add(u32 kind, u32 op, W a, W b)
    {
        Node n;
        n.kind = kind;
        n.op = op;
        n.a = a;
        n.b = b;
        u32 la = c->nodes[a].level, lb = kind == 1 ? c->nodes[b].level : 0;
        n.level = kind == 1 ? std::max(la, lb) + 1 : la;
        c->nodes.push_back(n);
        c->scheduled = false;
        if (kind == 1) {
            c->n_pbs++;
            c->depth = std::max(c->depth, n.level);
        }
        return (W)(c->nodes.size() - 1);
    }
    W nand_(W a, W b) { return add(1, PZ_NAND, a, b); }
    W and_(W a, W b) { return add(1, PZ_AND, a, b); }
    W or_(W a, W b) { return add(1, PZ_OR, a, b); }
    W nor_(W a, W b) { return add(1, PZ_NOR, a, b); }
    W xor_(W a, W b) { return add(1, PZ_XOR, a, b); }
    W xnor_(W a, W b) { return add(1, PZ_XNOR, a, b); }
    W not_(W a) { return add(2, GATE_NOT, a, a); }

    // half_adder, src/shortint/ops.rs:5-14
    W half_adder(W &a, W b)
    {
        W carry = and_(a, b);
        a = xor_(a, b);
        return carry;
    }
    // full_adder_plain_carry_in, :16-44
    W full_adder_plain(W &a, W b, bool carry_in)
    {
        W a_and_b = and_(a, b);
        a = xor_(a, b);
        if (carry_in) a_and_b = or_(a_and_b, a);
        if (carry_in) a = not_(a);
        return a_and_b;
    }
    // full_adder, :46-58
    W full_adder(W &a, W b, W carry_in)
    {
        W a_and_b = and_(a, b);
        a = xor_(a, b);
        W t = and_(a, carry_in);
        a_and_b = or_(a_and_b, t);
        a = xor_(a, carry_in);
        return a_and_b;
    }
    // arbitrary_bit_adder, :60-90 -> (carry_last, carry_last_last)
    std::pair<W, W> adder(std::vector<W> &a, const std::vector<W> &b, bool carry_in)
    {
        size_t n = a.size();
        W carry = !carry_in ? half_adder(a[0], b[0]) : full_adder_plain(a[0], b[0], true);
        for (size_t i = 1; i < 1 + (n - 3); i++) carry = full_adder(a[i], b[i], carry);
        W cll = full_adder(a[n - 2], b[n - 2], carry);
        W cl = full_adder(a[n - 1], b[n - 1], cll);
        return {cl, cll};
    }
    // arbitrary_bit_subtractor, :92-103 -> (difference, carry_last, carry_last_last)
    std::vector<W> subtractor(const std::vector<W> &a, const std::vector<W> &b, W *cl, W *cll)
    {
        std::vector<W> neg_b;
        for (W v : b) neg_b.push_back(not_(v));
        auto r = adder(neg_b, a, true);
        if (cl) *cl = r.first;
        if (cll) *cll = r.second;
        return neg_b;
    }
    // arbitrary_bit_mux, :119-137
    std::vector<W> mux(W sel, const std::vector<W> &t, const std::vector<W> &f)
    {
        W ns = not_(sel);
        std::vector<W> out;
        for (size_t i = 0; i < t.size(); i++) {
            W sa = and_(sel, t[i]);
            W sb = and_(ns, f[i]);
            sa = or_(sa, sb);
            out.push_back(sa);
        }
        return out;
    }
    // eight_bit_mul, :139-187
    std::vector<W> mul8(const std::vector<W> &a, const std::vector<W> &b)
    {
        std::vector<W> carries, out;
        for (int i = 0; i < 8; i++) {
            if (i == 0) {
                out.push_back(and_(a[0], b[0]));
            } else if (i == 1) {
                W tmp0 = and_(a[1], b[0]);
                W tmp1 = and_(a[0], b[1]);
                W carry = half_adder(tmp0, tmp1);
                carries.push_back(carry);
                out.push_back(tmp0);
            } else {
                W sum = and_(a[i], b[0]);
                W tmp = and_(a[i - 1], b[1]);
                carries[0] = full_adder(sum, tmp, carries[0]);
                for (int j = 2; j < i; j++) {
                    W t2 = and_(a[i - j], b[j]);
                    carries[j - 1] = full_adder(sum, t2, carries[j - 1]);
                }
                W t3 = and_(a[0], b[i]);
                W carry = half_adder(sum, t3);
                carries.push_back(carry);
                out.push_back(sum);
            }
        }
        return out;
    }
    // arbitrary_bit_division_for_quotient_and_rem, :189-257 -> (quotient, remainder)
    void divrem(const std::vector<W> &a, const std::vector<W> &b, std::vector<W> &quo,
                std::vector<W> &rem)
    {
        size_t n = a.size();
        std::vector<W> neg_b;
        for (W v : b) neg_b.push_back(not_(v));
        std::vector<W> remainder, quotient;
        for (size_t i = 0; i < n; i++) {
            remainder.push_back(a[n - 1 - i]);
            std::vector<W> subtract = remainder;
            W carry = full_adder_plain(subtract[i], neg_b[0], true);
            for (size_t j = 1; j < i + 1; j++) carry = full_adder(subtract[i - j], neg_b[j], carry);
            for (size_t j = i + 1; j < n; j++) carry = and_(carry, neg_b[j]);
            W not_carry = not_(carry);
            for (size_t k = 0; k < remainder.size(); k++) {
                subtract[k] = and_(subtract[k], carry);
                remainder[k] = and_(remainder[k], not_carry);
                remainder[k] = or_(remainder[k], subtract[k]);
            }
            quotient.push_back(carry);
        }
        std::reverse(remainder.begin(), remainder.end());
        std::reverse(quotient.begin(), quotient.end());
        quo = quotient;
        rem = remainder;
    }
    // is_zero, :259-270
    W is_zero(const std::vector<W> &a)
    {
        std::vector<W> na;
        for (W v : a) na.push_back(not_(v));
        W out = na[0];
        for (size_t i = 1; i < na.size(); i++) out = and_(out, na[i]);
        return out;
    }
    // arbitrary_bit_equality, :272-285
    W equal(const std::vector<W> &a, const std::vector<W> &b)
    {
        W out = xnor_(a[0], b[0]);
        for (size_t i = 1; i < a.size(); i++) {
            W e = xnor_(a[i], b[i]);
            out = and_(out, e);
        }
        return out;
    }
    // _comparator_handler_from_second_msb + arbitrary_bit_comparator, :287-356
    W comparator(const std::vector<W> &a, const std::vector<W> &b)
    {
        size_t n = a.size();
        W comp = not_(b[n - 1]);
        comp = and_(comp, a[n - 1]);
        W casc = xnor_(a[n - 1], b[n - 1]);
        W tmp = not_(b[n - 2]);
        tmp = and_(tmp, a[n - 2]);
        tmp = and_(tmp, casc);
        comp = or_(comp, tmp);
        for (size_t i = 2; i < n; i++) {
            W tc = xnor_(a[n - i], b[n - i]);
            casc = and_(casc, tc);
            W t2 = not_(b[n - 1 - i]);
            t2 = and_(t2, a[n - 1 - i]);
            t2 = and_(t2, casc);
            comp = or_(comp, t2);
        }
        return comp;
    }
    // set_div_by_zero_flag, src/shortint/mod.rs:42-56
    void div_flag(const std::vector<W> &den)
    {
        W z = is_zero(den);
        c->div_flag = c->div_flag < 0 ? (int)z : (int)or_((W)c->div_flag, z);
    }
};

// ctx may be NULL: the circuit can then be built and inspected (stats,
// nodes, level sizes) but not run.
extern "C" int pz_circuit_create(pz_ctx *ctx, pz_circuit **out)
{
    if (!out) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !out");
    pz_circuit *c = new pz_circuit();
    c->ctx = ctx;
    *out = c;
    return PZ_OK;
}

static int bind_ctx(pz_circuit *c)
{
    if (!c->ctx) return PZ_FAIL(PZ_ERR_STATE, "invalid state: !c->ctx");
    if (c->N1) return PZ_OK;
    pz_exec_view v;
    DevCtx dc;
    int rc = pz_internal_exec_view(c->ctx, 1, &v, &dc);
    if (rc) return rc;
    c->N1 = v.p.rlwe_n + 1;
    return PZ_OK;
}

extern "C" int pz_circuit_destroy(pz_circuit *c)
{
    if (!c) return PZ_OK;
    if (c->pool) (void)hipFree(c->pool);
    if (c->d_plan) (void)hipFree(c->d_plan);
    delete c;
    return PZ_OK;
}

extern "C" int pz_circuit_inputs(pz_circuit *c, uint32_t count, uint32_t *wires)
{
    if (!c || !wires) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c || !wires");
    for (uint32_t i = 0; i < count; i++) {
        Node n;
        n.kind = 0;
        n.op = 0;
        n.a = n.b = 0;
        n.level = 0;
        c->nodes.push_back(n);
        wires[i] = (uint32_t)(c->nodes.size() - 1);
        c->inputs.push_back(wires[i]);
    }
    c->scheduled = false;
    return PZ_OK;
}

static bool valid(pz_circuit *c, uint32_t w) { return w < c->nodes.size(); }

extern "C" int pz_circuit_gate(pz_circuit *c, uint32_t op, uint32_t a, uint32_t b, uint32_t *out)
{
    if (!c || !out || op > GATE_NOT || !valid(c, a) || (op != GATE_NOT && !valid(c, b)))
        return PZ_FAIL(PZ_ERR_ARG, "null argument, unknown gate or operand wire out of range");
    Builder bl{c};
    *out = op == GATE_NOT ? bl.not_(a) : bl.add(1, op, a, b);
    return PZ_OK;
}

extern "C" int pz_circuit_u8(pz_circuit *c, uint32_t op, const uint32_t *a8, const uint32_t *b8,
                             const uint32_t *sel, uint32_t *out, uint32_t *nout)
{
    if (!c || !a8 || !out || !nout)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c || !a8 || !out || !nout");
    if (!b8) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !b8");
    std::vector<W> a(a8, a8 + 8), b(b8, b8 + 8);
    for (size_t i = 0; i < 8; i++)
        if (!valid(c, a[i]) || !valid(c, b[i]))
            return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !valid(c, a[i]) || !valid(c, b[i])");
    Builder bl{c};
    std::vector<W> o;
    switch (op) {
    case PZ_U8_ADD: {
        o = a;
        bl.adder(o, b, false);
        break;
    }
    case PZ_U8_OVERFLOWING_ADD: {
        o = a;
        auto r = bl.adder(o, b, false);
        o.push_back(r.first);
        break;
    }
    case PZ_U8_SUB: o = bl.subtractor(a, b, nullptr, nullptr); break;
    case PZ_U8_OVERFLOWING_SUB: {
        W cl;
        o = bl.subtractor(a, b, &cl, nullptr);
        o.push_back(bl.not_(cl));
        break;
    }
    case PZ_U8_MUL: o = bl.mul8(a, b); break;
    case PZ_U8_DIV: {
        bl.div_flag(b);
        std::vector<W> qv, rv;
        bl.divrem(a, b, qv, rv);
        o = qv;
        break;
    }
    case PZ_U8_REM: {
        std::vector<W> qv, rv;
        bl.divrem(a, b, qv, rv);
        o = rv;
        break;
    }
    case PZ_U8_DIV_REM: {
        bl.div_flag(b);
        std::vector<W> qv, rv;
        bl.divrem(a, b, qv, rv);
        o = qv;
        o.insert(o.end(), rv.begin(), rv.end());
        break;
    }
    case PZ_U8_EQ: o = {bl.equal(a, b)}; break;
    case PZ_U8_NEQ: o = {bl.not_(bl.equal(a, b))}; break;
    case PZ_U8_LT: o = {bl.comparator(b, a)}; break;
    case PZ_U8_GT: o = {bl.comparator(a, b)}; break;
    case PZ_U8_LE: o = {bl.not_(bl.comparator(a, b))}; break;
    case PZ_U8_GE: o = {bl.not_(bl.comparator(b, a))}; break;
    case PZ_U8_MUX:
        if (!sel || !valid(c, *sel))
            return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !sel || !valid(c, *sel)");
        o = bl.mux(*sel, a, b);
        break;
    case PZ_U8_MAX: {
        W gt = bl.comparator(a, b);
        o = bl.mux(gt, a, b);
        break;
    }
    case PZ_U8_MIN: {
        W lt = bl.comparator(b, a);
        o = bl.mux(lt, a, b);
        break;
    }
    default: return PZ_FAIL(PZ_ERR_ARG, "unknown FheUint8 operation");
    }
    for (size_t i = 0; i < o.size(); i++) out[i] = o[i];
    *nout = (uint32_t)o.size();
    return PZ_OK;
}

extern "C" int pz_circuit_div_zero_flag(pz_circuit *c, uint32_t *wire)
{
    if (!c || !wire) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c || !wire");
    if (c->div_flag < 0) return PZ_FAIL(PZ_ERR_STATE, "invalid state: c->div_flag < 0");
    *wire = (uint32_t)c->div_flag;
    return PZ_OK;
}

extern "C" int pz_circuit_reset_flags(pz_circuit *c)
{
    if (!c) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c");
    c->div_flag = -1;
    return PZ_OK;
}

// Level schedule.  The circuit's depth D (its longest chain of bootstrapped
// gates) fixes the number of levels; gates off that chain have slack.  Each
// gate gets its latest level that still keeps depth D (ALAP), and levels are
// filled in order from the gates whose inputs are done, least slack first:
// every gate whose latest level has come, then others up to a width of whole
// rounds of the resident blind-rotation grid (at least the average of what
// remains per level).  Wide early levels are thereby spread into the late
// levels that a long chain (the running div-by-zero flag) leaves nearly
// empty.  NOT gates run in the batch of the gate that produces their input
// (level 0 for an input wire).  Order does not change any ciphertext: every
// gate reads the same inputs.
static void schedule(pz_circuit *c)
{
    if (c->scheduled) return;
    c->plan.clear();
    c->lplan.clear();
    c->plan_on_device = false;
    const size_t nn = c->nodes.size();
    const u32 D = c->depth;
    u32 cap = 512;  // bootstraps of one resident round (default without a device)
    if (c->ctx) {
        uint32_t cu = 0, per = 0;
        if (pz_ctx_info(c->ctx, nullptr, &cu, &per) == PZ_OK && cu * per > 0) cap = cu * per;
    }
    // producer of each wire's value: the bootstrapped gate behind any NOTs, or
    // none (an input wire)
    const u32 NONE = 0xffffffffu;
    std::vector<u32> src(nn, NONE);
    for (size_t i = 0; i < nn; i++) {
        const Node &n = c->nodes[i];
        src[i] = n.kind == 1 ? (u32)i : (n.kind == 2 ? src[n.a] : NONE);
    }
    // latest levels (ALAP) at depth D; nodes are in topological order
    std::vector<u32> alap(nn, D);
    for (size_t i = nn; i-- > 0;) {
        const Node &n = c->nodes[i];
        if (n.kind == 0) continue;
        const u32 lim = n.kind == 1 ? alap[i] - 1 : alap[i];
        alap[n.a] = std::min(alap[n.a], lim);
        if (n.kind == 1) alap[n.b] = std::min(alap[n.b], lim);
    }
    // bootstrapped-gate dependency graph
    std::vector<u32> indeg(nn, 0), level(nn, 0);
    std::vector<std::vector<u32>> users(nn);
    size_t remaining = 0;
    for (size_t i = 0; i < nn; i++) {
        const Node &n = c->nodes[i];
        if (n.kind != 1) continue;
        remaining++;
        const u32 pa = src[n.a], pb = src[n.b];
        if (pa != NONE) {
            indeg[i]++;
            users[pa].push_back((u32)i);
        }
        if (pb != NONE && pb != pa) {
            indeg[i]++;
            users[pb].push_back((u32)i);
        }
    }
    typedef std::pair<u32, u32> Item;  // (latest level, gate)
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> ready;
    for (size_t i = 0; i < nn; i++)
        if (c->nodes[i].kind == 1 && indeg[i] == 0) ready.push({alap[i], (u32)i});
    u32 nlev = 0;
    for (u32 L = 1; remaining > 0 && !ready.empty(); L++) {
        const size_t left = L <= D ? D - L + 1 : 1;
        const size_t avg = (remaining + left - 1) / left;
        const size_t width = (size_t)cap * ((std::max<size_t>(avg, 1) + cap - 1) / cap);
        std::vector<u32> take;
        while (!ready.empty() && (take.size() < width || ready.top().first <= L)) {
            take.push_back(ready.top().second);
            ready.pop();
        }
        for (u32 g : take) level[g] = L;
        for (u32 g : take)
            for (u32 u : users[g])
                if (--indeg[u] == 0) ready.push({alap[u], u});
        remaining -= take.size();
        nlev = L;
    }
    c->levels.assign((size_t)std::max(nlev, D) + 1, {});
    for (u32 i = 0; i < nn; i++) {
        const Node &n = c->nodes[i];
        if (n.kind == 0) continue;
        GateRec r;
        r.op = n.op;
        r.in0 = n.a;
        r.in1 = n.b;
        r.out = i;
        r.key = 0;
        const u32 l = n.kind == 1 ? level[i] : (src[i] == NONE ? 0 : level[src[i]]);
        c->levels[l].push_back(r);
    }
    c->scheduled = true;
}

// Level of every node in the schedule (input wires 0; a NOT runs in the
// batch of its input's producer, after its bootstrapped gates).
extern "C" int pz_circuit_node_levels(pz_circuit *c, uint32_t *h_levels, size_t cap)
{
    if (!c || !h_levels) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c || !h_levels");
    schedule(c);
    std::vector<uint32_t> lv(c->nodes.size(), 0);
    for (size_t l = 0; l < c->levels.size(); l++)
        for (auto &r : c->levels[l]) lv[r.out] = (uint32_t)l;
    for (size_t i = 0; i < lv.size() && i < cap; i++) h_levels[i] = lv[i];
    return PZ_OK;
}

extern "C" int pz_circuit_stats(pz_circuit *c, uint32_t *n_wires, uint32_t *n_pbs, uint32_t *depth)
{
    if (!c) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c");
    if (n_wires) *n_wires = (uint32_t)c->nodes.size();
    if (n_pbs) *n_pbs = c->n_pbs;
    if (depth) *depth = c->depth;
    return PZ_OK;
}

extern "C" int pz_circuit_level_sizes(pz_circuit *c, uint32_t *h_sizes, uint32_t cap)
{
    if (!c || !h_sizes) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c || !h_sizes");
    schedule(c);
    for (uint32_t l = 0; l < cap && l < c->levels.size(); l++) {
        uint32_t k = 0;
        for (auto &r : c->levels[l]) k += r.op <= 5;
        h_sizes[l] = k;
    }
    return PZ_OK;
}

static int ensure_pool(pz_circuit *c)
{
    size_t rows = c->nodes.size();
    if (rows <= c->pool_rows) return PZ_OK;
    if (c->pool) (void)hipFree(c->pool);
    c->pool = nullptr;
    size_t cap = std::max<size_t>(rows, 64);
    if (hipMalloc(&c->pool, 8 * cap * c->N1) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
    c->pool_rows = cap;
    return PZ_OK;
}

// The schedule's launch plan (host side): a NOT whose input an earlier NOT of
// its level writes runs in a later wave.
static void build_plan(pz_circuit *c)
{
    if (!c->lplan.empty() || c->levels.empty()) return;
    c->max_B = 0;
    for (auto &lvl : c->levels) {
        pz_circuit::LevelPlan lp;
        lp.off = c->plan.size();
        std::vector<const GateRec *> pbs, nots;
        for (auto &r : lvl) (r.op <= 5 ? pbs : nots).push_back(&r);
        lp.B = (u32)pbs.size();
        c->max_B = std::max(c->max_B, lp.B);
        for (auto *r : pbs) c->plan.push_back(r->op);
        for (auto *r : pbs) c->plan.push_back(r->in0);
        for (auto *r : pbs) c->plan.push_back(r->in1);
        for (auto *r : pbs) c->plan.push_back(r->out);
        std::map<u32, int> wave_of;  // NOT output wire -> its wave
        std::vector<std::vector<const GateRec *>> waves;
        for (auto *r : nots) {
            auto it = wave_of.find(r->in0);
            const int w = it == wave_of.end() ? 0 : it->second + 1;
            wave_of[r->out] = w;
            if ((int)waves.size() <= w) waves.resize(w + 1);
            waves[w].push_back(r);
        }
        for (auto &wv : waves) {
            pz_circuit::Wave w;
            w.off = c->plan.size();
            w.count = (u32)wv.size();
            for (auto *r : wv) c->plan.push_back(r->in0);
            for (auto *r : wv) c->plan.push_back(r->out);
            lp.waves.push_back(w);
        }
        c->lplan.push_back(lp);
    }
}

extern "C" int pz_circuit_run(pz_circuit *c, pz_key *key, const uint64_t *d_inputs, void *stream)
{
    if (!c || !key || (!d_inputs && !c->inputs.empty()))
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c || !key || (!d_inputs && !c->inputs.empty())");
    if (!pz_internal_key_fits(c->ctx, key))
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !pz_internal_key_fits(c->ctx, key)");
    int rc = bind_ctx(c);
    if (rc) return rc;
    schedule(c);
    rc = ensure_pool(c);
    if (rc) return rc;
    void *sv = nullptr;
    pz_ctx_stream(c->ctx, &sv);
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)sv;
    // inputs into their pool rows
    size_t ni = c->inputs.size();
    if (ni) {
        u32 *d_rows = nullptr;
        if (hipMalloc(&d_rows, 4 * ni) != hipSuccess)
            return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
        (void)hipMemcpyAsync(d_rows, c->inputs.data(), 4 * ni, hipMemcpyHostToDevice, st);
        hipLaunchKernelGGL(k_copy_rows, grid_rows(ni * c->N1), 256, 0, st, c->pool, d_inputs,
                           d_rows, ni, c->N1);
        (void)hipStreamSynchronize(st);
        (void)hipFree(d_rows);
    }
    // every level from one uploaded plan, launched back to back on the stream
    // (no host work or synchronisation between levels)
    build_plan(c);
    if (!c->plan_on_device && !c->plan.empty()) {
        // a new schedule: the previous one's launches (this stream, which does
        // not order with the null stream) may still read the old plan
        if (hipStreamSynchronize(st) != hipSuccess)
            return PZ_FAIL(PZ_ERR_HIP, "hipStreamSynchronize failed");
        if (c->d_plan_cap < c->plan.size()) {
            if (c->d_plan) (void)hipFree(c->d_plan);
            c->d_plan = nullptr;
            c->d_plan_cap = 0;
            if (hipMalloc(&c->d_plan, 4 * c->plan.size()) != hipSuccess)
                return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
            c->d_plan_cap = c->plan.size();
        }
        if (hipMemcpy(c->d_plan, c->plan.data(), 4 * c->plan.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
            return PZ_FAIL(PZ_ERR_HIP, "hipMemcpy failed");
        c->plan_on_device = true;
    }
    pz_exec_view v;
    DevCtx dc;
    rc = pz_internal_exec_view(c->ctx, c->max_B ? c->max_B : 1, &v, &dc);
    if (rc) return rc;
    const DevKey *k0 = pz_internal_devkey(key);
    const u32 N1 = v.p.rlwe_n + 1, n1 = v.p.lwe_n + 1;
    for (auto &lp : c->lplan) {
        const u32 *d = c->d_plan + lp.off;
        const size_t B = lp.B;
        if (B) {
            const bool rm = ks_rowmajor(*k0);  // the MFMA key switch's digit layout
            launch_prep(PREP_GATE, c->pool, c->pool, d + B, d + 2 * B, d, N1, B, v.digits, v.ksb,
                        v.ks, dc, st, rm);
            launch_ks(v.digits, v.ksb, k0->ksk, k0->ksk_bytes, v.p.rlwe_n * v.p.lwe_d, n1, k0->ldk,
                      0, B, v.ks, st, rm ? k0->ksk_planes : nullptr);
            int var = BR_ONE, grid = 1;
            pz_internal_br_plan(c->ctx, B, &var, &grid);
            launch_blind_rotate(v.logn, var, grid, v.ks, d, B, c->pool, v.stats, dc, *k0, nullptr,
                                nullptr, d + 3 * B, st);
        }
        for (auto &w : lp.waves)
            hipLaunchKernelGGL(k_not_rows, grid_rows((size_t)w.count * N1), 256, 0, st, c->pool,
                               c->d_plan + w.off, c->d_plan + w.off + w.count, (size_t)w.count, N1,
                               v.p.rlwe_q);
    }
    rc = PZ_LAUNCH_STATUS();
    if (rc == PZ_OK) c->ran_nodes = c->nodes.size();
    return rc;
}

extern "C" int pz_circuit_read(pz_circuit *c, const uint32_t *wires, size_t count, uint64_t *d_out,
                               void *stream)
{
    if (!c || !wires || !d_out) return PZ_FAIL(PZ_ERR_ARG, "null argument");
    for (size_t i = 0; i < count; i++)
        if (!valid(c, wires[i])) return PZ_FAIL(PZ_ERR_ARG, "wire out of range");
    if (!c->pool || c->ran_nodes == 0 || bind_ctx(c))
        return PZ_FAIL(PZ_ERR_STATE, "the circuit has not been run on a device");
    // the pool holds only the wires of the last run: a wire added since then
    // has no value yet (and may lie past the pool's rows)
    for (size_t i = 0; i < count; i++)
        if (wires[i] >= c->ran_nodes)
            return PZ_FAIL(PZ_ERR_STATE, "wire added after the last run: run the circuit again");
    void *sv = nullptr;
    pz_ctx_stream(c->ctx, &sv);
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)sv;
    u32 *d_rows = nullptr;
    if (hipMalloc(&d_rows, 4 * std::max<size_t>(count, 1)) != hipSuccess)
        return PZ_FAIL(PZ_ERR_HIP, "hipMalloc failed");
    (void)hipMemcpyAsync(d_rows, wires, 4 * count, hipMemcpyHostToDevice, st);
    hipLaunchKernelGGL(k_gather_rows, grid_rows(count * c->N1), 256, 0, st, c->pool, d_out, d_rows,
                       count, c->N1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_rows);
    return PZ_OK;
}

// Node list for inspection: kind, op, a, b per node (4 words each).
extern "C" int pz_circuit_nodes(pz_circuit *c, uint32_t *h_out, size_t cap)
{
    if (!c || !h_out) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !c || !h_out");
    for (size_t i = 0; i < c->nodes.size() && i < cap; i++) {
        h_out[4 * i] = c->nodes[i].kind;
        h_out[4 * i + 1] = c->nodes[i].op;
        h_out[4 * i + 2] = c->nodes[i].a;
        h_out[4 * i + 3] = c->nodes[i].b;
    }
    return PZ_OK;
}
