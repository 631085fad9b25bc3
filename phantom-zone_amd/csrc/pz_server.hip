// pz_server.hip -- serving loop for FheUint8 programs from many sessions.
//
// Sessions own server keys; a session's key is uploaded to HBM when the
// session is created if the keys' total footprint fits the budget, otherwise
// when one of its requests is admitted (idle resident sessions are then
// evicted, least recently used first).  Requests are gate DAGs (pz_circuit)
// with input ciphertexts and arrival times; they may be submitted from
// another thread while the loop runs, and then arrive when submitted.  Each
// loop step admits arrived requests as HBM and pool rows allow, then takes
// the next ready level of every active request -- across sessions, up to a
// bootstrap budget per launch -- and runs them as ONE batch (per-ciphertext
// keys), so requests that arrive while others are on the GPU join the next
// step (continuous batching).  The loop does not wait for a step to finish
// before queueing the next one (each context keeps up to two steps in
// flight); a request is complete when the event recorded after its last
// level has passed, and only then are its rows reused.  A server may hold sessions of several
// parameter sets (one context each, same device and ring size, one wire
// pool): a step forms one batch per context and launches them on the
// contexts' own streams, where they run concurrently.  The reference has no
// server; its evaluation is one gate at a time on one thread
// (src/bool/evaluator.rs:2136-2323).
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
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
    // requests: a deque keeps every element in place while submissions from
    // another thread append during a run; mu guards the appends, the request
    // states the other entry points read, the intake flag and the clock
    std::deque<Request> reqs;
    std::mutex mu;
    std::condition_variable cv;  // a submission wakes an idle loop
    bool intake_open = false;
    bool running = false;
    bool failed = false;  // a run ended in an error: request states are not kept
    std::chrono::steady_clock::time_point t0;
    std::vector<hipEvent_t> free_events;
    u64 *pool = nullptr;
    RowAlloc rows;
    u32 N1 = 0;
    size_t resident_bytes = 0;
    uint64_t tick = 0;
    // stats
    uint64_t total_pbs = 0, batches = 0, key_uploads = 0, evictions = 0;
    double wall = 0;
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
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
    for (auto e : s->free_events) (void)hipEventDestroy(e);
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
    if (!s || !prog || !out_wires || !rid) return PZ_FAIL(PZ_ERR_ARG, "null argument");
    if (session >= s->sessions.size()) return PZ_FAIL(PZ_ERR_ARG, "no such session");
    uint32_t nw = 0, npbs = 0, depth = 0;
    pz_circuit_stats(prog, &nw, &npbs, &depth);
    if (nw > s->cfg.pool_rows)
        return PZ_FAIL(PZ_ERR_UNSUPPORTED, "program has more wires than the pool has rows");
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
    if (!r.inputs.empty() && !h_inputs) return PZ_FAIL(PZ_ERR_ARG, "null input ciphertexts");
    r.h_in.assign(h_inputs, h_inputs + r.inputs.size() * (size_t)s->N1);
    for (uint32_t i = 0; i < n_out; i++) {
        if (out_wires[i] >= nw) return PZ_FAIL(PZ_ERR_ARG, "output wire out of range");
        r.outs.push_back(out_wires[i]);
    }
    r.n_ops = n_ops ? n_ops : 1;
    r.n_pbs = npbs;
    r.rows = nw;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        // a negative arrival is "now": the loop's clock while it runs, 0 before
        r.arrival = arrival >= 0 ? arrival : (s->running ? s->now() : 0.0);
        s->reqs.push_back(std::move(r));
        *rid = (uint32_t)(s->reqs.size() - 1);
    }
    s->cv.notify_one();
    return PZ_OK;
}

extern "C" int pz_server_intake(pz_server *s, int open)
{
    if (!s) return PZ_FAIL(PZ_ERR_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->intake_open = open != 0;
    }
    s->cv.notify_one();
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

static hipEvent_t take_event(pz_server *s)
{
    if (!s->free_events.empty()) {
        hipEvent_t e = s->free_events.back();
        s->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

extern "C" int pz_server_run(pz_server *s)
{
    if (!s) return PZ_FAIL(PZ_ERR_ARG, "null server");
    const size_t nc = s->ctxs.size();
    std::vector<hipStream_t> streams(nc);
    for (size_t c = 0; c < nc; c++) {
        void *sv = nullptr;
        pz_ctx_stream(s->ctxs[c], &sv);
        streams[c] = (hipStream_t)sv;
    }
    size_t scanned = 0, total = 0, done = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->failed) return PZ_FAIL(PZ_ERR_STATE, "an earlier run of this server failed");
        s->t0 = std::chrono::steady_clock::now();
        s->running = true;
        for (auto &r : s->reqs)
            if (r.state == 3) done++;
    }
    struct Finishing {
        size_t req;
        hipEvent_t ev;  // recorded after the request's last level
    };
    std::vector<size_t> waiting;  // taken in, arrival still ahead (sorted by arrival)
    std::deque<size_t> pending;   // arrived, not admitted
    std::vector<size_t> active;   // admitted, levels left to queue
    std::vector<Finishing> finishing;
    // the loop's view of the requests: element pointers taken under the lock
    // (a deque keeps its elements in place on push_back, but indexing it
    // while another thread appends reads a block map that may be replaced)
    std::vector<Request *> R;
    int rc = PZ_OK;
    for (;;) {
        bool open;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (; scanned < s->reqs.size(); scanned++) {
                R.push_back(&s->reqs[scanned]);
                if (R.back()->state == 0) waiting.push_back(scanned);
            }
            total = s->reqs.size();
            open = s->intake_open;
        }
        std::stable_sort(waiting.begin(), waiting.end(), [&](size_t x, size_t y) {
            return R[x]->arrival < R[y]->arrival;
        });
        double t = s->now();
        size_t k = 0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            while (k < waiting.size() && R[waiting[k]]->arrival <= t) {
                R[waiting[k]]->state = 1;
                pending.push_back(waiting[k++]);
            }
        }
        waiting.erase(waiting.begin(), waiting.begin() + k);
        // completions: outputs copied once the last level's event has passed,
        // then the rows go back to the pool
        for (size_t i = 0; i < finishing.size();) {
            const hipError_t q = hipEventQuery(finishing[i].ev);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            if (q != hipSuccess) {
                rc = PZ_FAIL(PZ_ERR_HIP, "a served batch failed");
                break;
            }
            Request &r = *R[finishing[i].req];
            std::vector<uint64_t> out(r.outs.size() * (size_t)s->N1);
            for (size_t o = 0; o < r.outs.size() && rc == PZ_OK; o++)
                if (hipMemcpy(out.data() + o * s->N1, s->pool + (r.base + r.outs[o]) * s->N1,
                              8 * s->N1, hipMemcpyDeviceToHost) != hipSuccess)
                    rc = PZ_FAIL(PZ_ERR_HIP, "hipMemcpy of a result failed");
            if (rc) break;
            s->rows.release(r.base, r.rows);
            s->sessions[r.session].active--;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                r.result.swap(out);
                r.finished = s->now();
                r.state = 3;
            }
            done++;
            s->free_events.push_back(finishing[i].ev);
            finishing.erase(finishing.begin() + i);
        }
        if (rc) break;
        // admission in arrival order (a request whose session cannot be made
        // resident does not block later requests of resident sessions)
        for (auto it = pending.begin(); it != pending.end();) {
            Request &r = *R[*it];
            if (active.size() + finishing.size() >= s->cfg.max_active) break;
            size_t base;
            if (!make_resident(s, r.session) || !s->rows.alloc(r.rows, &base)) {
                ++it;
                continue;
            }
            r.base = base;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                r.state = 2;
                r.admitted = s->now();
            }
            s->sessions[r.session].active++;
            // inputs into their rows: the rows have no work in flight, and
            // the synchronous copy completes before the request's first level
            // is queued
            for (size_t i = 0; i < r.inputs.size() && rc == PZ_OK; i++)
                if (hipMemcpy(s->pool + (r.base + r.inputs[i]) * s->N1,
                              r.h_in.data() + i * s->N1, 8 * s->N1,
                              hipMemcpyHostToDevice) != hipSuccess)
                    rc = PZ_FAIL(PZ_ERR_HIP, "hipMemcpy of an input failed");
            active.push_back(*it);
            it = pending.erase(it);
        }
        if (rc) break;
        // one batch per context: the next level of every active request of
        // its sessions, oldest first
        std::vector<std::vector<const pz_key *>> keys(nc);
        std::vector<int> key_of(s->sessions.size(), -1);  // index in its context's list
        std::vector<std::vector<GateRec>> recs(nc);
        std::vector<size_t> stepped, pbs(nc, 0);
        for (size_t idx : active) {
            Request &r = *R[idx];
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
        // queue every context's batch
        for (size_t c = 0; c < nc && rc == PZ_OK; c++) {
            if (recs[c].empty()) continue;
            rc = exec_batch(s->ctxs[c], keys[c], recs[c], s->pool, streams[c]);
            if (rc) break;
            s->total_pbs += pbs[c];
            s->batches++;
        }
        if (rc) break;
        s->tick++;
        for (size_t idx : stepped) {
            Request &r = *R[idx];
            s->sessions[r.session].last_used = s->tick;
            if (r.next_level < r.levels.size()) r.next_level++;
            while (r.next_level < r.levels.size() && r.levels[r.next_level].empty()) r.next_level++;
            if (r.next_level < r.levels.size()) continue;
            // every level queued: complete once its context's stream passes
            // an event recorded here
            const u32 c = s->sessions[r.session].ctx;
            hipEvent_t e = take_event(s);
            if (!e || hipEventRecord(e, streams[c]) != hipSuccess) {
                rc = PZ_FAIL(PZ_ERR_HIP, "hipEventRecord failed");
                break;
            }
            active.erase(std::find(active.begin(), active.end(), idx));
            finishing.push_back({idx, e});
        }
        if (rc) break;
        if (done == total && !open) break;
        if (!stepped.empty()) continue;
        // nothing queued this round: wait for a completion, an arrival or a
        // submission
        if (!finishing.empty()) {
            if (hipEventSynchronize(finishing.front().ev) != hipSuccess) {
                rc = PZ_FAIL(PZ_ERR_HIP, "a served batch failed");
                break;
            }
            continue;
        }
        if (!pending.empty() && waiting.empty() && !open) {
            rc = PZ_FAIL(PZ_ERR_STATE, "a request cannot be admitted (key budget or pool rows)");
            break;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->reqs.size() > scanned || s->intake_open != open) continue;
        double wait = 0.01;  // an open intake, or requests that cannot be admitted yet
        if (!waiting.empty())
            wait = std::min(wait, std::max(0.0, R[waiting.front()]->arrival - s->now()));
        if (wait > 0)
            s->cv.wait_for(lk, std::chrono::duration<double>(wait),
                           [&]() { return s->reqs.size() > scanned || s->intake_open != open; });
    }
    // on an error, let queued work finish before the events go back
    for (auto &f : finishing) {
        (void)hipEventSynchronize(f.ev);
        s->free_events.push_back(f.ev);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->running = false;
    s->failed = rc != PZ_OK;
    s->wall = s->now();
    return rc;
}

extern "C" int pz_server_result(pz_server *s, uint32_t rid, uint64_t *h_out)
{
    if (!s || !h_out) return PZ_FAIL(PZ_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (rid >= s->reqs.size()) return PZ_FAIL(PZ_ERR_ARG, "no such request");
    Request &r = s->reqs[rid];
    if (r.state != 3) return PZ_FAIL(PZ_ERR_STATE, "request not complete");
    memcpy(h_out, r.result.data(), 8 * r.result.size());
    return PZ_OK;
}

extern "C" int pz_server_request_times(pz_server *s, uint32_t rid, double *arrival,
                                       double *admitted, double *finished, uint32_t *n_pbs,
                                       uint32_t *n_ops)
{
    if (!s) return PZ_FAIL(PZ_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (rid >= s->reqs.size()) return PZ_FAIL(PZ_ERR_ARG, "no such request");
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
    if (!s) return PZ_FAIL(PZ_ERR_ARG, "null server");
    if (total_pbs) *total_pbs = s->total_pbs;
    if (batches) *batches = s->batches;
    if (key_uploads) *key_uploads = s->key_uploads;
    if (evictions) *evictions = s->evictions;
    if (wall_s) *wall_s = s->wall;
    return PZ_OK;
}
