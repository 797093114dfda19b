// Area queries (tools/dflood.cpp) and the request server on top of the C ABI (include/dsx.h).
//
// Server threads:
//   * the gate thread takes scan requests in arrival order, runs the GPU gate over their seed range in
//     chunks on its own context and cuts the gated seeds into groups;
//   * each worker thread owns a context and repeatedly takes the oldest pending work: either one group
//     of a scan (run_group: tables, scan, sites, hot zones, proxy and exact floods) or a batch of every
//     pending area query whose worst-case flood tiles fit its tile pool.
// A scan finishes when its gate is done and all its groups are; its rows are formatted in group
// order, so they equal those of the same scan run alone.
#include "../../include/dsx.h"
#include "../csrc/dsx_rng.h"
#include "pipeline.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace dsx {
extern thread_local std::string g_last_error;
}

namespace {

// Contexts of destroyed servers, kept for the next server on the same device (DSX_SERVER_POOL=0: none
// kept). Creating and destroying a server's 17 contexts (a stream, the flood engine's buffers, pinned
// staging each) one server after another in one process made later servers' bursts stall now and then;
// a context keeps no request state between calls, so a pooled one serves the next server as a new one
// would. dsx_server_pool_clear frees them. The pool is never destroyed: its contexts may outlive the
// runtime's static teardown.
struct CtxPool {
    std::mutex mu;
    std::map<int, std::vector<dsx_ctx *>> free;
    std::map<int, uint64_t> footprint;   // one worker's measured device bytes, per device
};
CtxPool &ctx_pool() {
    static CtxPool *p = new CtxPool;
    return *p;
}
bool pool_on() {
    static const bool on = [] {
        const char *e = std::getenv("DSX_SERVER_POOL");
        return !(e && *e == '0');
    }();
    return on;
}
// *fresh (optional): whether the context was created now rather than taken from the pool
int pool_get(int device, dsx_ctx **out, bool *fresh = nullptr) {
    if (fresh) *fresh = false;
    if (pool_on()) {
        CtxPool &P = ctx_pool();
        std::lock_guard<std::mutex> l(P.mu);
        auto &v = P.free[device];
        if (!v.empty()) {
            *out = v.back();
            v.pop_back();
            return DSX_OK;
        }
    }
    if (fresh) *fresh = true;
    return dsx_create(device, out);
}
void pool_put(int device, dsx_ctx *c) {
    if (!c) return;
    if (!pool_on()) {
        dsx_destroy(c);
        return;
    }
    CtxPool &P = ctx_pool();
    std::lock_guard<std::mutex> l(P.mu);
    P.free[device].push_back(c);
}

constexpr uint64_t TILE_CELLS = 32 * 32;
constexpr uint64_t POOL_TILES = 1ull << 21;   // tiles of a context's flood engine (floods_api.hip)

// worst-case tiles of a step-1 window of half-size h
uint64_t window_tiles(int h) {
    const uint64_t t = (uint64_t(2 * h + 1) + 31) / 32;
    return t * t;
}

}  // namespace

extern "C" {

int dsx_area_queries(dsx_ctx *ctx, const dsx_area_query *q, uint64_t n, dsx_area *out) {
    if (!ctx || (n && (!q || !out))) return DSX_E_ARG;
    if (n == 0) return DSX_OK;
    for (uint64_t k = 0; k < n; k++)
        if (q[k].half < 0 || q[k].half > 20000) return DSX_E_ARG;
    // start test: desert_state at the query quart (dflood.cpp:17)
    std::vector<uint64_t> seeds;
    std::unordered_map<uint64_t, uint32_t> slot;
    std::vector<uint32_t> sidx(n);
    std::vector<int32_t> xs(n), zs(n);
    for (uint64_t k = 0; k < n; k++) {
        auto [it, fresh] = slot.try_emplace(q[k].seed, uint32_t(seeds.size()));
        if (fresh) seeds.push_back(q[k].seed);
        sidx[k] = it->second;
        xs[k] = q[k].x >> 2;
        zs[k] = q[k].z >> 2;
    }
    std::vector<int8_t> st(n);
    int rc = dsx_points(ctx, seeds.data(), seeds.size(), sidx.data(), xs.data(), zs.data(), n, DSX_PRED_DESERT_STATE, 0.f,
                        st.data(), nullptr);
    if (rc) return rc;
    std::vector<dsx_flood_job> jobs;
    std::vector<uint64_t> who;
    for (uint64_t k = 0; k < n; k++) {
        out[k] = dsx_area{0, 1, xs[k], xs[k], zs[k], zs[k], 0, 0};
        if (st[k] != 2) continue;
        dsx_flood_job j;
        std::memset(&j, 0, sizeof j);
        j.seed = q[k].seed;
        j.kind = DSX_PRED_DESERT_STATE;
        j.qx = xs[k];
        j.qz = zs[k];
        j.step = 1;
        j.half = q[k].half;
        jobs.push_back(j);
        who.push_back(k);
    }
    // batches whose worst-case tiles fit the pool (a single window always does)
    size_t lo = 0;
    while (lo < jobs.size()) {
        size_t hi = lo;
        uint64_t tiles = 0;
        while (hi < jobs.size() && (hi == lo || tiles + window_tiles(jobs[hi].half) <= POOL_TILES) && hi - lo < 4096) {
            tiles += window_tiles(jobs[hi].half);
            hi++;
        }
        std::vector<dsx_area> res(hi - lo);
        if ((rc = dsx_flood(ctx, jobs.data() + lo, hi - lo, res.data()))) return rc;
        for (size_t k = lo; k < hi; k++) out[who[k]] = res[k - lo];
        lo = hi;
    }
    return DSX_OK;
}

void dsx_area_workload(uint64_t gen_seed, uint64_t n, dsx_area_query *out, uint64_t *caps) {
    dsx::Xoro r = dsx::Xoro::seeded(gen_seed);
    for (uint64_t j = 0; j < n; j++) {
        const int32_t X = int32_t(r.bounded(60000001u)) - 30000000;
        const int32_t Z = int32_t(r.bounded(60000001u)) - 30000000;
        const double u = r.uniform();
        const uint64_t cap = uint64_t(std::floor(std::pow(10.0, 3.0 + 4.0 * u)));
        dsx_area_query &o = out[j];
        std::memset(&o, 0, sizeof o);
        o.seed = dsx::mix64(j);
        o.x = X;
        o.z = Z;
        o.half = int32_t(std::floor((std::sqrt(double(cap)) - 1.0) / 2.0));
        if (caps) caps[j] = cap;
    }
}

}  // extern "C"

// ------------------------------------------------------------------ server
struct dsx_server {
    struct Req {
        uint64_t id = 0;
        int kind = 0;
        uint64_t start = 0, count = 0;      // scan
        dsx_area_query q{};                 // area
        int status = DSX_OK;
        std::string err;
        double t_submit = 0, t_start = 0, t_done = 0;
        bool done = false;
        dsx_area area{};
        dsx_search_counts counts{};
        // scan progress
        bool gate_done = false;
        size_t groups_total = 0, groups_done = 0;
        std::vector<std::unique_ptr<dsxh::Group>> groups;
        std::string csv, seeds;
    };
    struct GroupWork {
        Req *r;
        dsxh::Group *g;
        double t;   // arrival of the request (oldest-first order)
    };

    int device = 0;
    dsx_server_config cfg{};
    dsx_ctx *gate_ctx = nullptr;
    std::vector<dsx_ctx *> wctx;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_gate, cv_done;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_id = 1;
    std::deque<Req *> scan_q, area_q;
    std::deque<GroupWork> group_q;
    bool closing = false;
    std::vector<std::thread> threads;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    void finish_scan(Req *r) {   // with mu held: all groups done
        for (auto &g : r->groups) {
            if (g->rc && r->status == DSX_OK) {
                r->status = g->rc;
                r->err = g->err;
            }
            if (r->status) continue;
            dsxh::GroupText t;
            dsxh::format_group(*g, cfg.params, t, nullptr);
            r->csv += t.csv;
            r->seeds += t.seeds;
            r->counts.scanned += t.scanned;
            r->counts.sites += t.sites;
        }
        r->groups.clear();
        r->t_done = now();
        r->done = true;
        cv_done.notify_all();
    }

    void gate_loop() {
        for (;;) {
            Req *r;
            {
                std::unique_lock<std::mutex> l(mu);
                cv_gate.wait(l, [this] { return closing || !scan_q.empty(); });
                if (scan_q.empty()) return;
                r = scan_q.front();
                scan_q.pop_front();
                r->t_start = now();
            }
            const dsx_search_params &P = cfg.params;
            const dsx_gate_params gp{float(P.gate_phase), float(P.gate_amp), float(P.gate_pthr), 0.01f, 1.0f, 1};
            const uint64_t CHUNK = 1ull << 22;
            const size_t GROUP = 64;
            std::vector<uint64_t> batch;
            int rc = DSX_OK;
            std::string err;
            for (uint64_t pos = 0; pos < r->count && rc == DSX_OK;) {
                const uint64_t m = std::min(CHUNK, r->count - pos);
                uint64_t n = 0;
                dsx_gate_counts gc;
                batch.resize(1 << 16);
                rc = dsx_gate(gate_ctx, r->start + pos, m, &gp, batch.data(), nullptr, nullptr, batch.size(), &n, &gc);
                if (rc == DSX_E_CAPACITY) {
                    batch.resize(n);
                    rc = dsx_gate(gate_ctx, r->start + pos, m, &gp, batch.data(), nullptr, nullptr, batch.size(), &n, &gc);
                }
                if (rc) {
                    err = dsx_last_error();
                    break;
                }
                batch.resize(n);
                pos += m;
                std::lock_guard<std::mutex> l(mu);
                r->counts.done += n;
                r->counts.passed += n;
                for (size_t b0 = 0; b0 < batch.size(); b0 += GROUP) {
                    auto g = std::make_unique<dsxh::Group>();
                    g->seeds.assign(batch.begin() + b0, batch.begin() + std::min(batch.size(), b0 + GROUP));
                    group_q.push_back({r, g.get(), r->t_submit});
                    r->groups.push_back(std::move(g));
                    r->groups_total++;
                }
                cv_work.notify_all();
            }
            std::lock_guard<std::mutex> l(mu);
            if (rc && r->status == DSX_OK) {
                r->status = rc;
                r->err = err;
            }
            r->gate_done = true;
            if (r->groups_done == r->groups_total) finish_scan(r);
        }
    }

    void worker_loop(dsx_ctx *w) {
        // at most 64 queries per batch by default: a burst then spreads over every worker instead of filling
        // the first one's batch (4096 queries at once: p50 130 -> 35 ms, 21 k -> 23.5 k queries/s)
        const size_t max_batch = cfg.area_batch > 0 ? size_t(cfg.area_batch) : 64;
        for (;;) {
            std::vector<Req *> areas;
            GroupWork gw{nullptr, nullptr, 0};
            {
                std::unique_lock<std::mutex> l(mu);
                cv_work.wait(l, [this] { return closing || !area_q.empty() || !group_q.empty(); });
                if (area_q.empty() && group_q.empty()) return;
                const bool take_area = !area_q.empty() && (group_q.empty() || area_q.front()->t_submit <= group_q.front().t);
                if (take_area) {
                    uint64_t tiles = 0;
                    while (!area_q.empty() && areas.size() < max_batch) {
                        const uint64_t t = window_tiles(area_q.front()->q.half);
                        if (!areas.empty() && tiles + t > POOL_TILES) break;
                        tiles += t;
                        areas.push_back(area_q.front());
                        area_q.pop_front();
                    }
                    const double ts = now();
                    for (Req *r : areas) r->t_start = ts;
                } else {
                    gw = group_q.front();
                    group_q.pop_front();
                }
            }
            if (!areas.empty()) {
                std::vector<dsx_area_query> q(areas.size());
                std::vector<dsx_area> res(areas.size());
                for (size_t k = 0; k < areas.size(); k++) q[k] = areas[k]->q;
                const int rc = dsx_area_queries(w, q.data(), q.size(), res.data());
                const std::string err = rc ? dsx_last_error() : "";
                std::lock_guard<std::mutex> l(mu);
                const double td = now();
                for (size_t k = 0; k < areas.size(); k++) {
                    Req *r = areas[k];
                    r->status = rc;
                    r->err = err;
                    if (!rc) r->area = res[k];
                    r->t_done = td;
                    r->done = true;
                }
                cv_done.notify_all();
            } else if (gw.g) {
                const int rc = dsxh::run_group(w, cfg.params, DSX_SEARCH_GATE, *gw.g);
                gw.g->rc = rc;
                gw.g->err = rc ? dsx_last_error() : "";
                std::lock_guard<std::mutex> l(mu);
                gw.g->done = true;
                gw.r->groups_done++;
                if (gw.r->gate_done && gw.r->groups_done == gw.r->groups_total) finish_scan(gw.r);
            }
        }
    }

    void fill(const Req &r, dsx_result *o) const {
        std::memset(o, 0, sizeof *o);
        o->id = r.id;
        o->kind = r.kind;
        o->status = r.status;
        o->t_submit = r.t_submit;
        o->t_start = r.t_start;
        o->t_done = r.t_done;
        o->area = r.area;
        o->counts = r.counts;
        o->csv_bytes = r.csv.size();
        o->seeds_bytes = r.seeds.size();
    }
};

extern "C" {

int dsx_server_create(int device, const dsx_server_config *cfg, dsx_server **out) {
    if (!cfg || !out || cfg->workers < 0 || cfg->workers > 64 || cfg->area_batch < 0) return DSX_E_ARG;
    *out = nullptr;
    auto s = std::make_unique<dsx_server>();
    s->device = device;
    s->cfg = *cfg;
    int rc = pool_get(device, &s->gate_ctx);
    if (rc) return rc;
    const int want = cfg->workers ? cfg->workers : 16;
    int limit = want;
    size_t free0 = 0, free1 = 0, total = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free0, &total) != hipSuccess) rc = DSX_E_HIP;
    for (int k = 0; k < limit && rc == DSX_OK; k++) {
        dsx_ctx *w = nullptr;
        bool fresh = false;
        if ((rc = pool_get(device, &w, &fresh))) break;
        s->wctx.push_back(w);
        if (k > 0) continue;
        if (!fresh) {
            // a pooled first worker: its buffers are in place, and the server that created it measured them
            uint64_t per = 0;
            {
                CtxPool &P = ctx_pool();
                std::lock_guard<std::mutex> l(P.mu);
                per = P.footprint[device];
            }
            if (cfg->hbm_budget && per) limit = int(std::clamp<uint64_t>(cfg->hbm_budget / per, 1, uint64_t(want)));
            continue;
        }
        // the first worker's device footprint (buffers are allocated on first use: touch them) sets
        // how many workers the HBM budget admits
        dsx_area_query q{};
        q.seed = 1;
        q.half = 1;
        dsx_area a;
        if ((rc = dsx_area_queries(w, &q, 1, &a))) break;
        const dsx_scan_params sp{0.25f, 0.55f, 0.25f, 0.41f, 2.0, 2.0};
        float ma, mb;
        uint64_t nc = 0;
        dsx_raw_cand rcand[1];
        const uint64_t seed1 = 1;
        rc = dsx_scan(w, &seed1, 1, &sp, &ma, &mb, rcand, 1, &nc);
        if (rc == DSX_E_CAPACITY) rc = DSX_OK;
        if (rc) break;
        if (hipMemGetInfo(&free1, &total) != hipSuccess) {
            rc = DSX_E_HIP;
            break;
        }
        const uint64_t per = free0 > free1 ? uint64_t(free0 - free1) : 1;
        {
            CtxPool &P = ctx_pool();
            std::lock_guard<std::mutex> l(P.mu);
            P.footprint[device] = per;
        }
        if (cfg->hbm_budget) limit = int(std::clamp<uint64_t>(cfg->hbm_budget / per, 1, uint64_t(want)));
    }
    if (rc) {
        for (dsx_ctx *w : s->wctx) dsx_destroy(w);
        dsx_destroy(s->gate_ctx);
        return rc;
    }
    s->cfg.workers = int(s->wctx.size());
    dsx_server *sp = s.get();
    sp->threads.emplace_back([sp] { sp->gate_loop(); });
    for (dsx_ctx *w : sp->wctx) sp->threads.emplace_back([sp, w] { sp->worker_loop(w); });
    *out = s.release();
    return DSX_OK;
}

int dsx_server_destroy(dsx_server *s) {
    if (!s) return DSX_E_ARG;
    {
        std::unique_lock<std::mutex> l(s->mu);
        // finish what is in flight
        s->cv_done.wait(l, [s] {
            for (auto &[id, r] : s->reqs)
                if (!r->done) return false;
            return true;
        });
        s->closing = true;
    }
    s->cv_work.notify_all();
    s->cv_gate.notify_all();
    for (auto &t : s->threads) t.join();
    for (dsx_ctx *w : s->wctx) pool_put(s->device, w);
    pool_put(s->device, s->gate_ctx);
    delete s;
    return DSX_OK;
}

int dsx_server_pool_clear(int device) {
    std::vector<dsx_ctx *> v;
    {
        CtxPool &P = ctx_pool();
        std::lock_guard<std::mutex> l(P.mu);
        v.swap(P.free[device]);
    }
    for (dsx_ctx *c : v) dsx_destroy(c);
    return DSX_OK;
}

int dsx_server_workers(const dsx_server *s) { return s ? int(s->wctx.size()) : 0; }

int dsx_submit_scan(dsx_server *s, uint64_t start, uint64_t count, uint64_t *id) {
    if (!s || !id) return DSX_E_ARG;
    auto r = std::make_unique<dsx_server::Req>();
    std::lock_guard<std::mutex> l(s->mu);
    if (s->closing) return DSX_E_ARG;
    r->id = s->next_id++;
    r->kind = 0;
    r->start = start;
    r->count = count;
    r->t_submit = s->now();
    *id = r->id;
    s->scan_q.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv_gate.notify_one();
    return DSX_OK;
}

int dsx_submit_area(dsx_server *s, const dsx_area_query *q, uint64_t *id) {
    if (!s || !q || !id || q->half < 0 || q->half > 20000) return DSX_E_ARG;
    auto r = std::make_unique<dsx_server::Req>();
    std::lock_guard<std::mutex> l(s->mu);
    if (s->closing) return DSX_E_ARG;
    r->id = s->next_id++;
    r->kind = 1;
    r->q = *q;
    r->t_submit = s->now();
    *id = r->id;
    s->area_q.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv_work.notify_one();
    return DSX_OK;
}

int dsx_wait(dsx_server *s, uint64_t id, dsx_result *o) {
    if (!s || !o) return DSX_E_ARG;
    std::unique_lock<std::mutex> l(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return DSX_E_ARG;
    dsx_server::Req *r = it->second.get();
    s->cv_done.wait(l, [r] { return r->done; });
    s->fill(*r, o);
    if (r->status) dsx::g_last_error = r->err;
    return DSX_OK;
}

int dsx_poll(dsx_server *s, uint64_t id, dsx_result *o) {
    if (!s || !o) return DSX_E_ARG;
    std::lock_guard<std::mutex> l(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return DSX_E_ARG;
    if (!it->second->done) return DSX_E_BUSY;
    s->fill(*it->second, o);
    return DSX_OK;
}

int dsx_result_text(dsx_server *s, uint64_t id, char *csv, uint64_t csv_cap, char *seeds, uint64_t seeds_cap) {
    if (!s) return DSX_E_ARG;
    std::lock_guard<std::mutex> l(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return DSX_E_ARG;
    const dsx_server::Req &r = *it->second;
    if (!r.done) return DSX_E_BUSY;
    if ((csv && csv_cap < r.csv.size()) || (seeds && seeds_cap < r.seeds.size())) return DSX_E_CAPACITY;
    if (csv) std::memcpy(csv, r.csv.data(), r.csv.size());
    if (seeds) std::memcpy(seeds, r.seeds.data(), r.seeds.size());
    return DSX_OK;
}

int dsx_release(dsx_server *s, uint64_t id) {
    if (!s) return DSX_E_ARG;
    std::lock_guard<std::mutex> l(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || !it->second->done) return DSX_E_ARG;
    s->reqs.erase(it);
    return DSX_OK;
}

}  // extern "C"
