// pz_server.hip -- serving loop for FheUint8 programs from many sessions.
//
// Sessions own server keys; a session's key is uploaded to HBM when the
// session is created if the keys' total footprint fits the budget, otherwise
// when one of its requests is admitted (idle resident sessions are then
// evicted, least recently used first).  Requests
// are gate DAGs (pz_circuit) with input ciphertexts and arrival times.  Each
// loop step admits arrived requests as HBM and pool rows allow, then takes the
// next ready level of every active request -- across sessions, up to a
// bootstrap budget per launch -- and runs them as ONE batch (per-ciphertext
// keys), so requests that arrive while others are on the GPU join the next
// step (continuous batching).  A server may hold sessions of several
// parameter sets (one context each, same device and ring size, one wire
// pool): a step forms one batch per context and launches them on the
// contexts' own streams, where they run concurrently.  The reference has no
// server; its evaluation is one gate at a time on one thread
// (src/bool/evaluator.rs:2136-2323).
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "../../include/pz.h"
#include "pz_exec.hpp"

using namespace pz;

// circuit internals (pz_circuit.hip)
extern "C" int pz_circuit_nodes(pz_circuit *c, uint32_t *h_out, size_t cap);
extern "C" int pz_circuit_stats(pz_circuit *c, uint32_t *n_wires, uint32_t *n_pbs,
                                uint32_t *depth);

namespace {

struct Session {
    const uint64_t *h_rgsw, *h_auto, *h_ksk;
    u32 ctx = 0;  // index into pz_server::ctxs
    pz_key *key = nullptr;
    size_t bytes = 0;
    int active = 0;
    uint64_t last_used = 0;
};

struct Request {
    uint32_t session;
    std::vector<uint32_t> kind, op, a, b, level;  // DAG nodes
    std::vector<uint32_t> inputs;                 // input node ids
    std::vector<std::vector<uint32_t>> levels;    // node ids per level
    std::vector<uint64_t> h_in;                   // input ciphertexts
    std::vector<uint32_t> outs;                   // output node ids
    std::vector<uint64_t> result;                 // output ciphertexts
    uint32_t n_ops = 1, n_pbs = 0;
    double arrival = 0, admitted = -1, finished = -1;
    size_t base = 0, rows = 0;  // pool rows [base, base + rows)
    size_t next_level = 0;
    int state = 0;  // 0 waiting, 1 pending, 2 active, 3 done
};

// first-fit allocator over pool rows
struct RowAlloc {
    std::vector<std::pair<size_t, size_t>> free_;  // (begin, len)
    void init(size_t n) { free_ = {{0, n}}; }
    bool alloc(size_t n, size_t *out)
    {
        for (size_t i = 0; i < free_.size(); i++) {
            if (free_[i].second >= n) {
                *out = free_[i].first;
                free_[i].first += n;
                free_[i].second -= n;
                if (free_[i].second == 0) free_.erase(free_.begin() + i);
                return true;
            }
        }
        return false;
    }
    void release(size_t b, size_t n)
    {
        free_.push_back({b, n});
        std::sort(free_.begin(), free_.end());
        std::vector<std::pair<size_t, size_t>> m;
        for (auto &f : free_) {
            if (!m.empty() && m.back().first + m.back().second == f.first) m.back().second += f.second;
            else m.push_back(f);
        }
        free_ = m;
    }
};

}  // namespace

struct pz_server {
    std::vector<pz_ctx *> ctxs;  // parameter sets served (ctxs[0]: the creating context)
    pz_server_config cfg;
    std::vector<Session> sessions;
    std::vector<Request> reqs;
    u64 *pool = nullptr;
    RowAlloc rows;
    u32 N1 = 0;
    size_t resident_bytes = 0;
    uint64_t tick = 0;
    // stats
    uint64_t total_pbs = 0, batches = 0, key_uploads = 0, evictions = 0;
    double wall = 0;
};

extern "C" int pz_server_create(pz_ctx *ctx, const pz_server_config *cfg, pz_server **out)
{
    if (!ctx || !cfg || !out) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !ctx || !cfg || !out");
    if (cfg->pool_rows == 0 || cfg->max_batch_pbs == 0)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: cfg->pool_rows == 0 || cfg->max_batch_pbs == 0");
    pz_server *s = new pz_server();
    s->ctxs.push_back(ctx);
    s->cfg = *cfg;
    pz_exec_view v;
    DevCtx dc;
    int rc = pz_internal_exec_view(ctx, 1, &v, &dc);
    if (rc) {
        delete s;
        return rc;
    }
    s->N1 = v.p.rlwe_n + 1;
    if (hipMalloc(&s->pool, 8 * (size_t)cfg->pool_rows * s->N1) != hipSuccess) {
        delete s;
        return PZ_FAIL(PZ_ERR_HIP, "hipMalloc of the wire pool failed");
    }
    s->rows.init(cfg->pool_rows);
    *out = s;
    return PZ_OK;
}

extern "C" int pz_server_destroy(pz_server *s)
{
    if (!s) return PZ_OK;
    for (auto &se : s->sessions)
        if (se.key) pz_key_free(se.key);
    if (s->pool) (void)hipFree(s->pool);
    delete s;
    return PZ_OK;
}

// Registers a session of context index ci; its key goes resident now while
// the resident keys fit the budget without evicting anyone, else on admission.
static int add_session_on(pz_server *s, u32 ci, const uint64_t *h_rgsw, const uint64_t *h_auto,
                          const uint64_t *h_ksk, uint32_t *sid)
{
    Session se;
    se.h_rgsw = h_rgsw;
    se.h_auto = h_auto;
    se.h_ksk = h_ksk;
    se.ctx = ci;
    s->sessions.push_back(se);
    *sid = (uint32_t)(s->sessions.size() - 1);
    pz_ctx *ctx = s->ctxs[ci];
    if (s->resident_bytes + pz_internal_key_footprint(ctx) <= s->cfg.hbm_budget_bytes) {
        Session &ns = s->sessions.back();
        if (pz_key_upload(ctx, h_rgsw, h_auto, h_ksk, &ns.key) != PZ_OK) {
            ns.key = nullptr;
            return PZ_OK;  // retried on admission
        }
        pz_key_bytes(ns.key, &ns.bytes);
        s->resident_bytes += ns.bytes;
        s->key_uploads++;
    }
    return PZ_OK;
}

extern "C" int pz_server_add_session(pz_server *s, const uint64_t *h_rgsw, const uint64_t *h_auto,
                                     const uint64_t *h_ksk, uint32_t *sid)
{
    if (!s || !h_rgsw || !h_auto || !h_ksk || !sid)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !s || !h_rgsw || !h_auto || !h_ksk || !sid");
    return add_session_on(s, 0, h_rgsw, h_auto, h_ksk, sid);
}

extern "C" int pz_server_add_session_ctx(pz_server *s, pz_ctx *ctx, const uint64_t *h_rgsw,
                                         const uint64_t *h_auto, const uint64_t *h_ksk,
                                         uint32_t *sid)
{
    if (!s || !ctx || !h_rgsw || !h_auto || !h_ksk || !sid)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !s || !ctx || !h_rgsw || !h_auto || !h_ksk || !sid");
    u32 ci = 0;
    while (ci < s->ctxs.size() && s->ctxs[ci] != ctx) ci++;
    if (ci == s->ctxs.size()) {
        // another parameter set: same device, same ring size (pool rows)
        pz_exec_view v;
        DevCtx dc;
        int rc = pz_internal_exec_view(ctx, 1, &v, &dc);
        if (rc) return rc;
        if (pz_internal_ctx_device(ctx) != pz_internal_ctx_device(s->ctxs[0]) ||
            v.p.rlwe_n + 1 != s->N1)
            return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: pz_internal_ctx_device(ctx) != pz_internal_ctx_device(s->ctxs[0]) || v.p.rlwe_n + 1 != s->N1");
        s->ctxs.push_back(ctx);
    }
    return add_session_on(s, ci, h_rgsw, h_auto, h_ksk, sid);
}

extern "C" int pz_server_submit(pz_server *s, uint32_t session, pz_circuit *prog,
                                const uint64_t *h_inputs, const uint32_t *out_wires,
                                uint32_t n_out, uint32_t n_ops, double arrival, uint32_t *rid)
{
    if (!s || !prog || !out_wires || !rid || session >= s->sessions.size())
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !s || !prog || !out_wires || !rid || session >= s->sessions.size()");
    uint32_t nw = 0, npbs = 0, depth = 0;
    pz_circuit_stats(prog, &nw, &npbs, &depth);
    std::vector<uint32_t> nodes(4 * (size_t)nw);
    pz_circuit_nodes(prog, nodes.data(), nw);
    Request r;
    r.session = session;
    r.kind.resize(nw);
    r.op.resize(nw);
    r.a.resize(nw);
    r.b.resize(nw);
    r.level.resize(nw);
    uint32_t maxl = 0;
    for (uint32_t i = 0; i < nw; i++) {
        r.kind[i] = nodes[4 * i];
        r.op[i] = nodes[4 * i + 1];
        r.a[i] = nodes[4 * i + 2];
        r.b[i] = nodes[4 * i + 3];
        uint32_t l = 0;
        if (r.kind[i] == 1) l = std::max(r.level[r.a[i]], r.level[r.b[i]]) + 1;
        else if (r.kind[i] == 2) l = r.level[r.a[i]];
        r.level[i] = l;
        maxl = std::max(maxl, l);
        if (r.kind[i] == 0) r.inputs.push_back(i);
    }
    r.levels.assign(maxl + 1, {});
    for (uint32_t i = 0; i < nw; i++)
        if (r.kind[i] != 0) r.levels[r.level[i]].push_back(i);
    if (!r.inputs.empty() && !h_inputs)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !r.inputs.empty() && !h_inputs");
    r.h_in.assign(h_inputs, h_inputs + r.inputs.size() * (size_t)s->N1);
    for (uint32_t i = 0; i < n_out; i++) {
        if (out_wires[i] >= nw) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: out_wires[i] >= nw");
        r.outs.push_back(out_wires[i]);
    }
    r.n_ops = n_ops ? n_ops : 1;
    r.n_pbs = npbs;
    r.arrival = arrival;
    r.rows = nw;
    if (nw > s->cfg.pool_rows)
        return PZ_FAIL(PZ_ERR_UNSUPPORTED, "unsupported: nw > s->cfg.pool_rows");
    s->reqs.push_back(std::move(r));
    *rid = (uint32_t)(s->reqs.size() - 1);
    return PZ_OK;
}

static bool make_resident(pz_server *s, uint32_t sid)
{
    Session &se = s->sessions[sid];
    if (se.key) return true;
    const size_t need = pz_internal_key_footprint(s->ctxs[se.ctx]);
    while (s->resident_bytes > 0 && s->resident_bytes + need > s->cfg.hbm_budget_bytes) {
        // evict the least recently used idle session
        int victim = -1;
        for (size_t i = 0; i < s->sessions.size(); i++) {
            Session &o = s->sessions[i];
            if (o.key && o.active == 0 && (victim < 0 || o.last_used < s->sessions[victim].last_used))
                victim = (int)i;
        }
        if (victim < 0) return false;
        Session &v = s->sessions[victim];
        pz_key_free(v.key);
        v.key = nullptr;
        s->resident_bytes -= v.bytes;
        s->evictions++;
    }
    if (pz_key_upload(s->ctxs[se.ctx], se.h_rgsw, se.h_auto, se.h_ksk, &se.key) != PZ_OK) {
        se.key = nullptr;
        return false;
    }
    pz_key_bytes(se.key, &se.bytes);
    if (s->resident_bytes + se.bytes > s->cfg.hbm_budget_bytes && s->resident_bytes > 0) {
        pz_key_free(se.key);
        se.key = nullptr;
        return false;
    }
    s->resident_bytes += se.bytes;
    s->key_uploads++;
    return true;
}

extern "C" int pz_server_run(pz_server *s)
{
    if (!s) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !s");
    const size_t nc = s->ctxs.size();
    std::vector<hipStream_t> streams(nc);
    for (size_t c = 0; c < nc; c++) {
        void *sv = nullptr;
        pz_ctx_stream(s->ctxs[c], &sv);
        streams[c] = (hipStream_t)sv;
    }
    auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    std::vector<size_t> order(s->reqs.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return s->reqs[x].arrival < s->reqs[y].arrival; });
    size_t next_arrival = 0, done = 0;
    std::deque<size_t> pending;
    std::vector<size_t> active;
    for (auto &r : s->reqs)
        if (r.state == 3) done++;
    while (done < s->reqs.size()) {
        double t = now();
        while (next_arrival < order.size() && s->reqs[order[next_arrival]].arrival <= t) {
            Request &r = s->reqs[order[next_arrival]];
            if (r.state == 0) {
                r.state = 1;
                pending.push_back(order[next_arrival]);
            }
            next_arrival++;
        }
        // admission in arrival order (a request whose session cannot be made
        // resident does not block later requests of resident sessions)
        for (auto it = pending.begin(); it != pending.end();) {
            Request &r = s->reqs[*it];
            if (active.size() >= s->cfg.max_active) break;
            size_t base;
            if (!make_resident(s, r.session) || !s->rows.alloc(r.rows, &base)) {
                ++it;
                continue;
            }
            r.base = base;
            r.state = 2;
            r.admitted = now();
            s->sessions[r.session].active++;
            // inputs into their pool rows, on the stream of the session's context
            for (size_t i = 0; i < r.inputs.size(); i++)
                (void)hipMemcpyAsync(s->pool + (r.base + r.inputs[i]) * s->N1,
                                     r.h_in.data() + i * s->N1, 8 * s->N1,
                                     hipMemcpyHostToDevice, streams[s->sessions[r.session].ctx]);
            active.push_back(*it);
            it = pending.erase(it);
        }
        if (active.empty()) {
            if (next_arrival < order.size()) {
                double wait = s->reqs[order[next_arrival]].arrival - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
                continue;
            }
            if (!pending.empty()) return PZ_FAIL(PZ_ERR_STATE, "invalid state: !pending.empty()");  // cannot admit (budget too small)
            break;
        }
        // one batch per context: the next level of every active request of
        // its sessions, oldest first
        std::vector<std::vector<const pz_key *>> keys(nc);
        std::vector<int> key_of(s->sessions.size(), -1);  // index in its context's list
        std::vector<std::vector<GateRec>> recs(nc);
        std::vector<size_t> stepped, pbs(nc, 0);
        for (size_t idx : active) {
            Request &r = s->reqs[idx];
            const u32 c = s->sessions[r.session].ctx;
            while (r.next_level < r.levels.size() && r.levels[r.next_level].empty()) r.next_level++;
            if (r.next_level >= r.levels.size()) {
                stepped.push_back(idx);
                continue;
            }
            size_t lp = 0;
            for (u32 n : r.levels[r.next_level]) lp += r.kind[n] == 1;
            if (pbs[c] > 0 && pbs[c] + lp > s->cfg.max_batch_pbs) continue;
            if (key_of[r.session] < 0) {
                key_of[r.session] = (int)keys[c].size();
                keys[c].push_back(s->sessions[r.session].key);
            }
            for (u32 n : r.levels[r.next_level]) {
                GateRec g;
                g.op = r.kind[n] == 2 ? (u32)GATE_NOT : r.op[n];
                g.in0 = (u32)(r.base + r.a[n]);
                g.in1 = (u32)(r.base + r.b[n]);
                g.out = (u32)(r.base + n);
                g.key = (u32)key_of[r.session];
                recs[c].push_back(g);
            }
            pbs[c] += lp;
            stepped.push_back(idx);
        }
        // launch every context's batch before waiting for any of them
        for (size_t c = 0; c < nc; c++) {
            if (recs[c].empty()) continue;
            int rc = exec_batch(s->ctxs[c], keys[c], recs[c], s->pool, streams[c]);
            if (rc) return rc;
        }
        for (size_t c = 0; c < nc; c++) {
            if (recs[c].empty()) continue;
            if (hipStreamSynchronize(streams[c]) != hipSuccess)
                return PZ_FAIL(PZ_ERR_HIP, "hipStreamSynchronize failed");
            s->total_pbs += pbs[c];
            s->batches++;
        }
        s->tick++;
        for (size_t idx : stepped) {
            Request &r = s->reqs[idx];
            s->sessions[r.session].last_used = s->tick;
            if (r.next_level < r.levels.size()) r.next_level++;
            while (r.next_level < r.levels.size() && r.levels[r.next_level].empty()) r.next_level++;
            if (r.next_level >= r.levels.size()) {
                // complete: fetch outputs, release rows
                r.result.resize(r.outs.size() * (size_t)s->N1);
                hipStream_t rs = streams[s->sessions[r.session].ctx];
                for (size_t i = 0; i < r.outs.size(); i++)
                    (void)hipMemcpyAsync(r.result.data() + i * s->N1,
                                         s->pool + (r.base + r.outs[i]) * s->N1, 8 * s->N1,
                                         hipMemcpyDeviceToHost, rs);
                (void)hipStreamSynchronize(rs);
                r.finished = now();
                r.state = 3;
                s->rows.release(r.base, r.rows);
                s->sessions[r.session].active--;
                done++;
                active.erase(std::find(active.begin(), active.end(), idx));
            }
        }
    }
    s->wall = now();
    return PZ_OK;
}

extern "C" int pz_server_result(pz_server *s, uint32_t rid, uint64_t *h_out)
{
    if (!s || rid >= s->reqs.size() || !h_out)
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !s || rid >= s->reqs.size() || !h_out");
    Request &r = s->reqs[rid];
    if (r.state != 3) return PZ_FAIL(PZ_ERR_STATE, "invalid state: r.state != 3");
    memcpy(h_out, r.result.data(), 8 * r.result.size());
    return PZ_OK;
}

extern "C" int pz_server_request_times(pz_server *s, uint32_t rid, double *arrival,
                                       double *admitted, double *finished, uint32_t *n_pbs,
                                       uint32_t *n_ops)
{
    if (!s || rid >= s->reqs.size())
        return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !s || rid >= s->reqs.size()");
    Request &r = s->reqs[rid];
    if (arrival) *arrival = r.arrival;
    if (admitted) *admitted = r.admitted;
    if (finished) *finished = r.finished;
    if (n_pbs) *n_pbs = r.n_pbs;
    if (n_ops) *n_ops = r.n_ops;
    return PZ_OK;
}

extern "C" int pz_server_stats(pz_server *s, uint64_t *total_pbs, uint64_t *batches,
                               uint64_t *key_uploads, uint64_t *evictions, double *wall_s)
{
    if (!s) return PZ_FAIL(PZ_ERR_ARG, "invalid argument: !s");
    if (total_pbs) *total_pbs = s->total_pbs;
    if (batches) *batches = s->batches;
    if (key_uploads) *key_uploads = s->key_uploads;
    if (evictions) *evictions = s->evictions;
    if (wall_s) *wall_s = s->wall;
    return PZ_OK;
}
