// The per-seed search of dsearch_desert.cpp:173-557 on top of the C ABI (include/dsx.h).
//
// Each seed runs as a C++20 coroutine that follows the reference worker's control flow statement by
// statement: sites -> hot zones -> proxy floods -> exact floods -> rows. Whenever it needs the GPU
// (points or floods) it co_awaits a request; the scheduler collects the requests of every seed in
// flight and executes them as one dsx_points / dsx_flood batch, then resumes the seeds. Where the
// reference's later steps depend only on coordinates (the hottest point of every site, the step-256
// pre-flood of every hot point, the proxy test of every start), the coroutine requests them all up
// front in one batch, so a seed needs only a handful of scheduler rounds. Order-dependent host state
// (hotcovered, covered, seen, std::sort of sites and proxies) is kept exactly as in the reference.
#include "../../include/dsx.h"
#include "../csrc/dsx_rng.h"
#include "../data/mc_tables.inc"
#include "cpu_ref.h"
#include "pipeline.h"
#include <atomic>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <coroutine>
#include <deque>
#include <mutex>
#include <thread>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <exception>
#include <map>
#include <optional>
#include <memory>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

namespace dsx {
extern thread_local std::string g_last_error;
}
int dsx_ctx_device(const dsx_ctx *c);
void dsx_ctx_free_workers(dsx_ctx *c);

namespace dsxh {
namespace cpu = dsx::cpu;

// gscancheck=N (dsearch_desert.cpp:228-242, 289-296): the first N seeds this process scans also get the
// reference's CPU tables and scan (cpu_ref.h); the line compares the CPU and GPU maxima and the buckets'
// best scores. gpucheck=N (:445-450): the first N big hot zones compare the GPU proxy predicate with the
// CPU DesertProxy at every 7th cell of their step-32 grids.
static std::atomic<int64_t> g_gscan_checked{0}, g_gpu_checked{0};

struct Site { float s, t, c, e; int x, z; };
struct FloodOut;
struct Eval {
    size_t rank;
    int x, z;
    double proxy, s8 = -1, s4 = 0;
    bool edge = false;
    double hot = 0;
    const FloodOut *px = nullptr;   // the proxy flood of this component (its counted cells hint the exact floods)
};

// ------------------------------------------------------------------ coroutine plumbing
struct Task {
    struct promise_type {
        Task get_return_object() { return Task{std::coroutine_handle<promise_type>::from_promise(*this)}; }
        std::suspend_never initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_void() {}
        void unhandled_exception() { std::terminate(); }
    };
    std::coroutine_handle<promise_type> h;
};

// A flood's outputs: the DenseArea and its counted cells (state 3 of dense_flood.hpp's window) as
// 32x32 tile masks sorted by (tz, tx).
struct FloodOut {
    dsx_area a{};
    int qx = 0, qz = 0, step = 1, half = 0;
    std::vector<dsx_tile_mask> tiles;
    bool counted(int i, int j) const {
        if (i < 0 || j < 0 || i > 2 * half || j > 2 * half) return false;
        const int tx = i >> 5, tz = j >> 5;
        auto it = std::lower_bound(tiles.begin(), tiles.end(), std::make_pair(tz, tx),
                                   [](const dsx_tile_mask &m, const std::pair<int, int> &k) {
                                       return m.tz != k.first ? m.tz < k.first : m.tx < k.second;
                                   });
        if (it == tiles.end() || it->tz != tz || it->tx != tx) return false;
        return (it->rows[j & 31] >> (i & 31)) & 1u;
    }
    // counted cells in window row-major order (j, then i), as the reference scans its bbox
    template <class F> void each_counted(F &&f) const {
        size_t t0 = 0;
        while (t0 < tiles.size()) {
            size_t t1 = t0;
            while (t1 < tiles.size() && tiles[t1].tz == tiles[t0].tz) t1++;
            for (int r = 0; r < 32; r++)
                for (size_t t = t0; t < t1; t++)
                    for (uint32_t b = tiles[t].rows[r]; b; b &= b - 1) f(tiles[t].tx * 32 + __builtin_ctz(b), tiles[t].tz * 32 + r);
            t0 = t1;
        }
    }
};

// The reference's key sets (hotcovered: 128-quart keys, covered: 32-quart keys) as the list of floods
// whose counted cells were inserted. A flood's keys are an affine image of its window lattice, so a
// key is in the set iff the cell it maps back to is counted in one of the floods.
struct KeySet {
    // kx0..kx1, kz0..kz1: the key range of the flood's counted cells (its bbox, plus the sub-points)
    struct Cover { const FloodOut *f; int shift, sub, bx, bz, kx0, kx1, kz0, kz1; };
    std::vector<Cover> covers;
    // Covers by 8192-quart cell of the offset quart plane: a cover is listed in every cell its key range
    // overlaps, so a query reads one cell's list instead of every cover (membership is an OR over the
    // covers, so the order they are tried in does not matter).
    static constexpr int CELL = 13;
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    static uint64_t cell_key(int64_t cx, int64_t cz) { return (uint64_t(uint32_t(cx)) << 32) | uint32_t(cz); }
    // insert every counted cell of f, lattice point x0 + (i - half) * step plus `sub` sub-points per axis
    // spaced 2^shift (sub = 2: the step-256 pre-flood's 4 keys per cell, dsearch_desert.cpp:345)
    void add(const FloodOut &f, int shift, int sub) {
        const Cover c{&f, shift, sub, (f.qx - f.half * f.step + 8000000) >> shift,
                      (f.qz - f.half * f.step + 8000000) >> shift, (f.a.minx + 8000000) >> shift,
                      ((f.a.maxx + 8000000) >> shift) + sub - 1, (f.a.minz + 8000000) >> shift,
                      ((f.a.maxz + 8000000) >> shift) + sub - 1};
        covers.push_back(c);
        if (f.a.cells == 0) return;   // never matches
        // offset quarts k with k >> shift in [kx0, kx1]: [kx0 << shift, ((kx1 + 1) << shift) - 1]
        const int64_t x0 = (int64_t(c.kx0) * (int64_t(1) << shift)) >> CELL,
                      x1 = ((int64_t(c.kx1) + 1) * (int64_t(1) << shift) - 1) >> CELL,
                      z0 = (int64_t(c.kz0) * (int64_t(1) << shift)) >> CELL,
                      z1 = ((int64_t(c.kz1) + 1) * (int64_t(1) << shift) - 1) >> CELL;
        for (int64_t gz = z0; gz <= z1; gz++)
            for (int64_t gx = x0; gx <= x1; gx++) grid[cell_key(gx, gz)].push_back(uint32_t(covers.size() - 1));
    }
    bool has(int x, int z) const {
        const int kx = (x + 8000000), kz = (z + 8000000);
        const auto it = grid.find(cell_key(int64_t(kx) >> CELL, int64_t(kz) >> CELL));
        if (it == grid.end()) return false;
        for (const uint32_t ci : it->second) {
            const Cover &c = covers[ci];
            if (c.f->a.cells == 0) continue;
            const int qx = kx >> c.shift, qz = kz >> c.shift;
            if (qx < c.kx0 || qx > c.kx1 || qz < c.kz0 || qz > c.kz1) continue;
            const int mx = qx - c.bx, mz = qz - c.bz;
            if (mx < 0 || mz < 0) continue;
            if (c.f->counted(mx / c.sub, mz / c.sub)) return true;
        }
        return false;
    }
};

// hash of a quart coordinate pair
struct XYHash {
    size_t operator()(const std::pair<int, int> &p) const {
        return std::hash<uint64_t>()((uint64_t(uint32_t(p.first)) << 32) | uint32_t(p.second));
    }
};

struct Sched;

struct PointsReq {
    Sched *s;
    uint64_t seed;
    int what;
    float cut;
    std::vector<int32_t> xs, zs;
    std::vector<double> vd;
    std::vector<int8_t> v8;
    std::coroutine_handle<> h;
    bool await_ready() const noexcept { return xs.empty(); }
    void await_suspend(std::coroutine_handle<> hh);
    void await_resume() const noexcept {}
};
struct FloodReq {
    Sched *s;
    std::vector<dsx_flood_job> jobs;
    bool want_cells;
    std::vector<FloodOut> out;
    std::coroutine_handle<> h;
    std::vector<dsx_flood_hint> hints;   // evaluation hints (job = index into jobs)
    bool await_ready() const noexcept { return jobs.empty(); }
    void await_suspend(std::coroutine_handle<> hh);
    void await_resume() const noexcept {}
};

// Wall-time accounting of the search stages, printed to stderr when DSX_PROFILE is set.
struct Prof {
    double t[20] = {};
    uint64_t n[20] = {};
    ~Prof() {
        if (!std::getenv("DSX_PROFILE")) return;
        static const char *name[20] = {"gate", "scan", "sites", "points", "floods", "rows", "steps", "flood jobs",
                                       "make_sites", "resume", "dsx_flood", "masks", "steps<=2", "steps<=8",
                                       "steps<=32", "steps>32", "zone starts", "proxy rounds", "site loop",
                                       "exact"};
        for (int i = 0; i < 20; i++)
            if (t[i] > 0 || n[i]) std::fprintf(stderr, "dsx_search profile: %-10s %8.3f s  %llu\n", name[i], t[i], (unsigned long long)n[i]);
        dsx_flood_stats f;
        dsx_flood_stats_get(&f, 0);
        std::fprintf(stderr, "dsx_flood profile: %llu calls, %llu rounds; small-window floods %llu (%llu sampled)\n",
                     (unsigned long long)f.calls, (unsigned long long)f.rounds, (unsigned long long)f.small_floods,
                     (unsigned long long)f.small_sampled);
        std::fprintf(stderr, "dsx_flood profile: %llu evaluation rounds, %.2f sub-blocks needed per round\n",
                     (unsigned long long)f.eval_rounds, f.eval_rounds ? double(f.eval_needed) / f.eval_rounds : 0.0);
        if (f.wg_life)
            std::fprintf(stderr, "dsx_flood profile: workgroup time %.3f s: waiting %.1f%%, evaluating %.1f%% (octave "
                         "loads %.1f%%), rest %.1f%%\n", f.wg_life * 1e-8, 100.0 * f.wg_wait / f.wg_life,
                         100.0 * f.wg_eval / f.wg_life, 100.0 * f.wg_load / f.wg_life,
                         100.0 * (double(f.wg_life) - f.wg_wait - f.wg_eval) / f.wg_life);
        static const char *kind[4] = {"desert_state", "hot", "proxy", "desert"};
        for (int k = 0; k < 4; k++)
            if (f.floods[k])
                std::fprintf(stderr, "dsx_flood profile: %-12s %8llu floods %9llu tiles, sampled %llu cold + %llu hot\n", kind[k],
                             (unsigned long long)f.floods[k], (unsigned long long)f.tiles[k],
                             (unsigned long long)f.sampled_cold[k], (unsigned long long)f.sampled_hot[k]);
    }
};
Prof g_prof;
std::mutex g_prof_mu;
std::atomic<int> g_interrupt{0};   // dsx_interrupt
void prof_add(int k, double t, uint64_t n) {
    std::lock_guard<std::mutex> l(g_prof_mu);
    g_prof.t[k] += t;
    g_prof.n[k] += n;
}
struct Timer {
    int k;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit Timer(int k_) : k(k_) {}
    ~Timer() { prof_add(k, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 1); }
};

struct Sched {
    dsx_ctx *ctx;
    std::vector<PointsReq *> points;
    std::vector<FloodReq *> floods;
    int err = DSX_OK;
    std::vector<uint32_t> moff, mpos, midx;   // flood_range's scratch: masks grouped by job

    int run_points(std::vector<PointsReq *> &rs) {
        // group by (what, cut); each group one dsx_points call over many seeds
        std::map<std::pair<int, float>, std::vector<PointsReq *>> groups;
        for (auto *r : rs) groups[{r->what, r->cut}].push_back(r);
        for (auto &[key, g] : groups) {
            const bool hottest = key.first == DSX_POINT_HOTTEST;
            const bool dbl = key.first == DSX_POINT_T || key.first == DSX_POINT_PROXYPROB || key.first == DSX_POINT_PROXY_C ||
                             key.first == DSX_POINT_PROXY_E || hottest;
            size_t i0 = 0;
            while (i0 < g.size()) {
                std::vector<uint64_t> seeds;
                std::unordered_map<uint64_t, uint32_t> slot;
                std::vector<uint32_t> sidx;
                std::vector<int32_t> xs, zs;
                size_t i1 = i0;
                for (; i1 < g.size(); i1++) {
                    PointsReq *r = g[i1];
                    if (!slot.count(r->seed)) {
                        if (seeds.size() >= 4000) break;
                        slot.emplace(r->seed, uint32_t(seeds.size()));
                        seeds.push_back(r->seed);
                    }
                    const uint32_t sl = slot[r->seed];
                    sidx.insert(sidx.end(), r->xs.size(), sl);
                    xs.insert(xs.end(), r->xs.begin(), r->xs.end());
                    zs.insert(zs.end(), r->zs.begin(), r->zs.end());
                }
                std::vector<double> vd(dbl ? xs.size() : 0);
                std::vector<int8_t> v8(dbl && !hottest ? 0 : xs.size());
                const int rc = dsx_points(ctx, seeds.data(), seeds.size(), sidx.data(), xs.data(), zs.data(), xs.size(),
                                          key.first, key.second, v8.data(), vd.data());
                if (rc) return rc;
                size_t off = 0;
                for (size_t i = i0; i < i1; i++) {
                    PointsReq *r = g[i];
                    const size_t m = r->xs.size();
                    if (dbl) r->vd.assign(vd.begin() + off, vd.begin() + off + m);
                    if (!dbl || hottest) r->v8.assign(v8.begin() + off, v8.begin() + off + m);
                    off += m;
                }
                i0 = i1;
            }
        }
        return DSX_OK;
    }

    // One dsx_flood call over jobs [lo, hi) of the flattened list; split in halves on capacity errors.
    int flood_range(std::vector<dsx_flood_job> &jobs, std::vector<std::pair<FloodReq *, size_t>> &who, size_t lo,
                    size_t hi, bool cells, const std::vector<dsx_flood_hint> &hints) {
        if (lo >= hi) return DSX_OK;
        std::vector<dsx_area> res(hi - lo);
        std::vector<dsx_flood_hint> sub;   // the hints of jobs [lo, hi), re-based
        for (const dsx_flood_hint &h : hints)
            if (h.job >= lo && h.job < hi) sub.push_back({uint32_t(h.job - lo), h.tx, h.tz, h.blocks});
        int rc;
        {
            Timer tf(10);
            rc = dsx_flood_hinted(ctx, jobs.data() + lo, hi - lo, sub.data(), sub.size(), res.data());
        }
        if (rc == DSX_E_CAPACITY && hi - lo > 1) {
            const size_t mid = lo + (hi - lo) / 2;
            if ((rc = flood_range(jobs, who, lo, mid, cells, hints))) return rc;
            return flood_range(jobs, who, mid, hi, cells, hints);
        }
        if (rc) return rc;
        for (size_t k = lo; k < hi; k++) {
            FloodOut &o = who[k].first->out[who[k].second];
            o.a = res[k - lo];
            o.qx = jobs[k].qx;
            o.qz = jobs[k].qz;
            o.step = jobs[k].step;
            o.half = jobs[k].half;
        }
        if (cells) {
            // each wanted job's masks, sorted by (tz, tx), read once from where the engine left them:
            // mask indices grouped by job (counting sort), each group's indices sorted by tile
            Timer tmk(11);
            const dsx_tile_mask *ma = nullptr, *mb = nullptr;
            uint64_t na = 0, nb = 0;
            if ((rc = flood_masks_view(ctx, &ma, &na, &mb, &nb))) return rc;
            auto mask = [&](uint32_t i) -> const dsx_tile_mask & { return i < na ? ma[i] : mb[i - na]; };
            const uint32_t nm = uint32_t(na + nb), nj = uint32_t(hi - lo);
            moff.assign(size_t(nj) + 1, 0);
            for (uint32_t i = 0; i < nm; i++) {
                const uint32_t j = mask(i).job;
                if (j >= nj) {
                    dsx::g_last_error = "flood mask of a job outside the batch";
                    return DSX_E_HIP;
                }
                moff[j + 1]++;
            }
            for (uint32_t k = 0; k < nj; k++) moff[k + 1] += moff[k];
            mpos.assign(moff.begin(), moff.end() - 1);
            midx.resize(nm);
            for (uint32_t i = 0; i < nm; i++) midx[mpos[mask(i).job]++] = i;
            for (uint32_t k = 0; k < nj; k++) {
                auto [r, q] = who[lo + k];
                if (!r->want_cells || moff[k + 1] == moff[k]) continue;
                const auto b = midx.begin() + moff[k], e = midx.begin() + moff[k + 1];
                if (e - b > 1)
                    std::sort(b, e, [&](uint32_t x, uint32_t y) {
                        const dsx_tile_mask &p = mask(x), &o = mask(y);
                        return p.tz != o.tz ? p.tz < o.tz : p.tx < o.tx;
                    });
                auto &t = r->out[q].tiles;
                t.resize(size_t(e - b));
                for (size_t u = 0; u < t.size(); u++) t[u] = mask(b[u]);
            }
        }
        return DSX_OK;
    }

    int run_floods(std::vector<FloodReq *> &rs) {
        std::vector<dsx_flood_job> jobs;
        std::vector<std::pair<FloodReq *, size_t>> who;
        std::vector<dsx_flood_hint> hints;
        bool cells = false;
        for (auto *r : rs) {
            r->out.assign(r->jobs.size(), FloodOut{});
            cells |= r->want_cells;
            for (const dsx_flood_hint &h : r->hints)
                hints.push_back({uint32_t(jobs.size() + h.job), h.tx, h.tz, h.blocks});
            for (size_t k = 0; k < r->jobs.size(); k++) {
                jobs.push_back(r->jobs[k]);
                who.push_back({r, k});
            }
        }
        // chunks: at most 4000 distinct seeds and 2^18 floods per call
        size_t lo = 0;
        while (lo < jobs.size()) {
            std::unordered_set<uint64_t> seen;
            size_t hi = lo;
            while (hi < jobs.size() && hi - lo < (1u << 18)) {
                if (!seen.count(jobs[hi].seed)) {
                    if (seen.size() >= 4000) break;
                    seen.insert(jobs[hi].seed);
                }
                hi++;
            }
            const int rc = flood_range(jobs, who, lo, hi, cells, hints);
            if (rc) return rc;
            lo = hi;
        }
        return DSX_OK;
    }

    // Execute all pending requests, then resume their coroutines. Returns false when nothing was
    // pending or a request failed (the caller then destroys the suspended coroutines).
    bool step() {
        if (err || (points.empty() && floods.empty())) return false;
        std::vector<PointsReq *> ps;
        std::vector<FloodReq *> fs;
        ps.swap(points);
        fs.swap(floods);
        prof_add(6, 0, 1);
        {
            const size_t nr = ps.size() + fs.size();
            prof_add(nr <= 2 ? 12 : nr <= 8 ? 13 : nr <= 32 ? 14 : 15, 0, 1);
        }
        if (!ps.empty()) {
            Timer t(3);
            err = run_points(ps);
        }
        if (!err && !fs.empty()) {
            Timer t(4);
            uint64_t nj = 0;
            for (auto *r : fs) nj += r->jobs.size();
            prof_add(7, 0, nj);
            err = run_floods(fs);
        }
        if (err) return false;
        Timer tr(9);
        for (auto *r : ps) r->h.resume();
        for (auto *r : fs) r->h.resume();
        return true;
    }
};
void PointsReq::await_suspend(std::coroutine_handle<> hh) {
    h = hh;
    s->points.push_back(this);
}
void FloodReq::await_suspend(std::coroutine_handle<> hh) {
    h = hh;
    s->floods.push_back(this);
}

// ------------------------------------------------------------------ per-seed work
struct SeedOut {
    uint64_t seed = 0;
    double fy[2] = {0, 0};
    float maxta = 0, maxtb = 0;
    bool scanned = false;
    size_t ncand = 0;
    std::vector<Site> sites;
    std::vector<Eval> ev;
    double best = 0;
    std::vector<std::string> lines;   // stderr lines (CANDIDATE)
    std::vector<std::pair<int, int>> cand_xy;   // block coordinates of each CANDIDATE line
    std::vector<dsx_area> cand_area;            // their exact step-1 areas (mcverify=1)
    std::vector<double> bests;        // per exact flood value v, in order (for the global "best" lines)
    std::vector<std::pair<int, int>> best_xy;
    std::vector<double> best_proxy;
    std::vector<int> best_step;
    std::vector<std::string> zones;   // PREFIX.zones.csv rows (study=1)
    uint64_t nhot = 0, nbig = 0;      // hot-zone floods and big zones (dsearch_desert.cpp:348, 369)
    bool done = false;
};

inline uint64_t hotkey(int x, int z) {
    return (uint64_t(uint32_t((x + 8000000) >> 7)) << 32) | uint32_t((z + 8000000) >> 7);
}
inline uint64_t cellkey(int x, int z) {
    return (uint64_t(uint32_t((x + 8000000) >> 5)) << 32) | uint32_t((z + 8000000) >> 5);
}
inline dsx_flood_job fjob(uint64_t seed, int kind, float cut, int qx, int qz, int step, int half, int conn8) {
    dsx_flood_job j;
    std::memset(&j, 0, sizeof j);
    j.seed = seed;
    j.kind = kind;
    j.cut = cut;
    j.qx = qx;
    j.qz = qz;
    j.step = step;
    j.half = half;
    j.conn8 = conn8;
    return j;
}

// Counted cells of a 4-connected flood in the order dense_flood_bridged (dense_flood.hpp:53-84) tests
// them: its depth-first traversal replayed with the flood's own mask as the predicate. A cell outside
// the mask that the traversal touches read 0 there, so the replay takes the same path.
void flood_order(const FloodOut &f, std::vector<std::pair<int, int>> &cells) {
    cells.clear();
    if (!f.a.cells) return;
    const int n = 2 * f.half + 1;
    thread_local std::vector<uint8_t> seen;
    thread_local std::vector<int> stack;
    seen.assign(size_t(n) * n, 0);
    stack.clear();
    auto touch = [&](int i, int j) {
        if (i < 0 || j < 0 || i >= n || j >= n) return;
        uint8_t &s = seen[size_t(j) * n + i];
        if (s) return;
        s = 1;
        if (!f.counted(i, j)) return;
        cells.push_back({f.qx + (i - f.half) * f.step, f.qz + (j - f.half) * f.step});
        stack.push_back(j * n + i);
    };
    touch(f.half, f.half);
    while (!stack.empty()) {
        const int k = stack.back();
        stack.pop_back();
        const int i = k % n, j = k / n;
        touch(i - 1, j);
        touch(i + 1, j);
        touch(i, j - 1);
        touch(i, j + 1);
    }
}

// study=1 (dsearch_desert.cpp:401-420): for every proxy component of the seed, in the order found, the
// hot zone (T >= 0.55 at step 64) flooded from the component's start, and the low-octave C/E census of
// its cells. study_points lists the counted cells of every zone (off: first cell of each zone);
// study_rows writes one PREFIX.zones.csv row per zone from their C and E values.
void study_points(const std::vector<FloodOut> &hz, PointsReq &q, std::vector<size_t> &off) {
    std::vector<std::pair<int, int>> cells;
    off.assign(hz.size() + 1, 0);
    for (size_t k = 0; k < hz.size(); k++) {
        flood_order(hz[k], cells);
        for (auto [x, z] : cells) {
            q.xs.push_back(x);
            q.zs.push_back(z);
        }
        off[k + 1] = q.xs.size();
    }
}
void study_rows(const dsx_search_params &P, SeedOut &O, const std::vector<FloodOut> &hz, const std::vector<double> &cv,
                const std::vector<double> &evv, const std::vector<size_t> &off) {
    const uint64_t seed = O.seed;
    const double phT = std::abs(O.fy[0] - .5) + std::abs(O.fy[1] - .5);
    const double phE0 = dsx::octave_phase(seed, DSX_CH_E, 0), phE1 = dsx::octave_phase(seed, DSX_CH_E, 1);
    const double phC0 = dsx::octave_phase(seed, DSX_CH_C, 0), phC1 = dsx::octave_phase(seed, DSX_CH_C, 1);
    for (size_t k = 0; k < hz.size(); k++) {
        double sc = 0, se = 0, se2 = 0;
        long nh = 0, no = 0, nb = 0, nm = 0, ng = 0;
        for (size_t p = off[k]; p < off[k + 1]; p++) {
            const double c = cv[p], e = evv[p];
            nh++;
            sc += c;
            se += e;
            se2 += e * e;
            no += c < -0.19;
            nb += e < 0.05;
            nm += e > 0.55;
            const int i = std::clamp(int(std::floor((c + 0.8) / 0.05)), 0, 39);
            const int j = std::clamp(int(std::floor((e + 1.0) / 0.05)), 0, 39);
            ng += DSX_DPROB[i][j] >= P.pcut;
        }
        const double d = double(std::max(1L, nh)), me = se / d;
        const Eval &z = O.ev[k];
        const Site &st = O.sites[z.rank];
        char line[512];
        std::snprintf(line, sizeof line,
                      "%lld,%.4f,%.4f,%.4f,%.4f,%.4f,%.4f,%zu,%d,%d,%.4f,%.4f,%.4f,%.2f,%.2f,%.4f,%.4f,%.4f,%.4f,%.4f,%.4f,%.4f\n",
                      (long long)seed, phT, O.maxta + O.maxtb, phE0, phE1, phC0, phC1, z.rank, z.x * 4, z.z * 4, st.t, st.c,
                      st.e, hz[k].a.cells * 64.0 * 64 * 16 / 1e6, z.proxy, double(no) / d, double(nb) / d, double(nm) / d,
                      double(ng) / d, sc / d, me, std::sqrt(std::max(0.0, se2 / d - me * me)));
        O.zones.push_back(line);
    }
}

// dsearch_desert.cpp:309-533 for one seed whose sites are known.
// Hints from a grid of marked sub-blocks (b, c) in [b0, b0 + w) x [c0, c0 + h): one hint per 32 x 32 tile
// (4 x 4 sub-blocks) holding a marked one, in tile order.
static void emit_hints(const std::vector<uint8_t> &mark, int b0, int c0, int w, int h, uint32_t job,
                       std::vector<dsx_flood_hint> &out) {
    const int t0x = b0 / 4, t0z = c0 / 4, ntx = (b0 + w - 1) / 4 - t0x + 1, ntz = (c0 + h - 1) / 4 - t0z + 1;
    std::vector<uint32_t> bits(size_t(ntx) * ntz, 0);
    for (int cc = 0; cc < h; cc++)
        for (int bb = 0; bb < w; bb++)
            if (mark[size_t(cc) * w + bb]) {
                const int b = b0 + bb, c = c0 + cc;
                bits[size_t(c / 4 - t0z) * ntx + (b / 4 - t0x)] |= 1u << ((c % 4) * 4 + b % 4);
            }
    for (int tz = 0; tz < ntz; tz++)
        for (int tx = 0; tx < ntx; tx++)
            if (const uint32_t m = bits[size_t(tz) * ntx + tx]) out.push_back({job, t0x + tx, t0z + tz, m});
}

// Evaluation hints for a step-4 flood that refines the step-8 flood f8 from the same centre
// (dsearch_desert.cpp:512-523): step-4 window cell (2i, 2j) is step-8 cell (i, j), so step-4 sub-block
// (b, c) (8 x 8 cells) covers step-8 cells [4b, 4b + 4) x [4c, 4c + 4). Every sub-block holding a counted
// step-8 cell, and its 8 neighbours, is hinted: the step-4 component lies within about one step-8 cell of
// the step-8 component, and a sub-block the hints miss is still evaluated when the flood touches it.
static void refine_hints(const FloodOut &f8, uint32_t job, int half4, std::vector<dsx_flood_hint> &out) {
    if (f8.tiles.empty()) return;
    int b0 = INT_MAX, b1 = INT_MIN, c0 = INT_MAX, c1 = INT_MIN;
    for (const dsx_tile_mask &m : f8.tiles) {
        b0 = std::min(b0, m.tx * 8);
        b1 = std::max(b1, m.tx * 8 + 7);
        c0 = std::min(c0, m.tz * 8);
        c1 = std::max(c1, m.tz * 8 + 7);
    }
    const int nb = (2 * half4 + 1 + 7) / 8;   // sub-blocks per axis of the step-4 window
    b0 = std::max(b0 - 1, 0), c0 = std::max(c0 - 1, 0);
    b1 = std::min(b1 + 1, nb - 1), c1 = std::min(c1 + 1, nb - 1);
    if (b0 > b1 || c0 > c1) return;
    const int w = b1 - b0 + 1, h = c1 - c0 + 1;
    std::vector<uint8_t> mark(size_t(w) * h, 0);
    for (const dsx_tile_mask &m : f8.tiles)
        for (int r = 0; r < 32; r++) {
            const uint32_t bits = m.rows[r];
            if (!bits) continue;
            const int c = m.tz * 8 + r / 4;
            for (int g = 0; g < 8; g++) {
                if (!((bits >> (4 * g)) & 0xfu)) continue;
                const int b = m.tx * 8 + g;
                for (int dc = -1; dc <= 1; dc++)
                    for (int db = -1; db <= 1; db++) {
                        const int bb = b + db, cc = c + dc;
                        if (bb >= b0 && bb <= b1 && cc >= c0 && cc <= c1) mark[size_t(cc - c0) * w + (bb - b0)] = 1;
                    }
            }
        }
    emit_hints(mark, b0, c0, w, h, job, out);
}

// Evaluation hints for a step-8 exact flood (window centre (sx, sz), half `half8`) from the proxy flood px
// whose component selected it (dsearch_desert.cpp:498-512): each counted proxy point stands for the 32 x 32
// quarts around it; every step-8 sub-block (8 x 8 cells, 64 x 64 quarts) within 16 + 8 quarts of a counted
// proxy point is hinted. The exact component mostly lies inside the proxy component; cells outside the hints
// are evaluated when the flood touches them.
static void proxy_hints(const FloodOut &px, uint32_t job, int sx, int sz, int half8, std::vector<dsx_flood_hint> &out) {
    if (px.tiles.empty() || px.a.cells == 0) return;
    const int n8 = 2 * half8 + 1, nb = (n8 + 7) / 8;
    auto fdiv = [](int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
    auto sub = [&](int q, int s0) { return fdiv(fdiv(q - s0, 8) + half8, 8); };   // step-8 sub-block of quart q
    // the sub-blocks within reach of the component's bounding box
    const int b0 = std::max(sub(px.a.minx - 24, sx), 0), b1 = std::min(sub(px.a.maxx + 24, sx), nb - 1);
    const int c0 = std::max(sub(px.a.minz - 24, sz), 0), c1 = std::min(sub(px.a.maxz + 24, sz), nb - 1);
    if (b0 > b1 || c0 > c1) return;
    const int w = b1 - b0 + 1, h = c1 - c0 + 1;
    std::vector<uint8_t> mark(size_t(w) * h, 0);
    px.each_counted([&](int i, int j) {
        const int x = px.qx + (i - px.half) * px.step, z = px.qz + (j - px.half) * px.step;
        for (int c = std::max(sub(z - 24, sz), c0); c <= std::min(sub(z + 24, sz), c1); c++)
            for (int b = std::max(sub(x - 24, sx), b0); b <= std::min(sub(x + 24, sx), b1); b++)
                mark[size_t(c - c0) * w + (b - b0)] = 1;
    });
    emit_hints(mark, b0, c0, w, h, job, out);
}

Task seed_task(Sched &S, const dsx_search_params &P, SeedOut &O) {
    const uint64_t seed = O.seed;
    const float pcut = float(P.pcut);
    const auto &sites = O.sites;
    const size_t ns = sites.size();
    std::vector<Eval> &ev = O.ev;
    std::vector<std::pair<int, int>> seen;
    KeySet covered, hotcovered;
    std::vector<std::unique_ptr<FloodOut>> keep;   // floods referenced by the key sets

    // ---- hottest of the 7x7 T samples of every site (dsearch_desert.cpp:328-332), one batch
    std::vector<int> hx(ns), hz(ns);
    std::vector<double> bt(ns, -9);
    // proxy-probability search of every site for the hot_min <= 0, multi <= 1 mode (:374-378)
    std::vector<int> px_x(ns), px_z(ns);
    std::vector<float> px_b(ns, -1);
    if (P.hot_min > 0) {
        // the device keeps the first maximum of the 7 x 7 T samples of each site (dz, then dx)
        PointsReq q{&S, seed, DSX_POINT_HOTTEST, 0.f, {}, {}, {}, {}, {}};
        q.xs.reserve(ns);
        q.zs.reserve(ns);
        for (size_t r = 0; r < ns; r++) {
            q.xs.push_back(sites[r].x);
            q.zs.push_back(sites[r].z);
        }
        co_await q;
        for (size_t r = 0; r < ns; r++) {
            const int k = q.v8[r];
            bt[r] = q.vd[r];
            hx[r] = sites[r].x + (k % 7 - 3) * 256;
            hz[r] = sites[r].z + (k / 7 - 3) * 256;
        }
    } else if (P.multi <= 1) {
        PointsReq q{&S, seed, DSX_POINT_PROXYPROB, pcut, {}, {}, {}, {}, {}};
        for (size_t r = 0; r < ns; r++)
            for (int dz = -4; dz <= 4; dz++)
                for (int dx = -4; dx <= 4; dx++) {
                    q.xs.push_back(sites[r].x + dx * 64);
                    q.zs.push_back(sites[r].z + dz * 64);
                }
        co_await q;
        size_t k = 0;
        for (size_t r = 0; r < ns; r++) {
            px_x[r] = sites[r].x;
            px_z[r] = sites[r].z;
            for (int dz = -4; dz <= 4; dz++)
                for (int dx = -4; dx <= 4; dx++, k++) {
                    const float pr = float(q.vd[k]);
                    if (pr > px_b[r]) {
                        px_b[r] = pr;
                        px_x[r] = sites[r].x + dx * 64;
                        px_z[r] = sites[r].z + dz * 64;
                    }
                }
        }
    }
    // ---- speculative batches. Every flood below is a pure function of its start point, so results
    // computed ahead are exactly what the sequential loop would compute; the loop decides which to use.
    using XY = std::pair<int, int>;
    std::map<XY, size_t> pre_idx, h128_idx;
    std::vector<size_t> pre_of(ns, 0), h128_of(ns, 0);   // per site: its hot point's pre-flood and step-128 flood
    FloodReq pre{&S, {}, true, {}, {}};          // step-256 pre-floods of every distinct hot point
    // A pre-flood only decides cells * 256^2 * 16 / 1e6 < hot_pre * hot_min (dsearch_desert.cpp:337-339)
    // and its cells are used only when it does; it may stop at the smallest count that fails the test.
    int pre_stop = 0;
    if (P.hot_pre > 0) {
        const double thr = P.hot_pre * P.hot_min;
        int64_t c = std::max<int64_t>(0, int64_t(thr / (256.0 * 256 * 16 / 1e6)) - 2);
        while (c * 256.0 * 256 * 16 / 1e6 < thr) c++;
        pre_stop = c >= 1 && c <= (1 << 23) ? int(c) : 0;
    }
    FloodReq h128{&S, {}, true, {}, {}};         // step-128 floods of every big one
    if (P.hot_min > 0) {
        for (size_t r = 0; r < ns; r++) {
            if (bt[r] < 0.55) continue;
            const auto [it, fresh] = pre_idx.try_emplace({hx[r], hz[r]}, pre.jobs.size());
            pre_of[r] = it->second;
            if (!fresh) continue;
            pre.jobs.push_back(fjob(seed, DSX_PRED_HOT, 0.f, hx[r], hz[r], 256, 32000 / 256, 0));
            pre.jobs.back().stop_cells = pre_stop;
        }
        if (P.hot_pre > 0) co_await pre;
        for (auto &[xy, k] : pre_idx)
            if (P.hot_pre <= 0 || pre.out[k].a.cells * 256.0 * 256 * 16 / 1e6 >= P.hot_pre * P.hot_min) {
                h128_idx[xy] = h128.jobs.size();
                h128.jobs.push_back(fjob(seed, DSX_PRED_HOT, 0.f, xy.first, xy.second, 128, 32000 / 128, 0));
            }
        for (size_t r = 0; r < ns; r++)
            if (bt[r] >= 0.55) {
                const auto it = h128_idx.find({hx[r], hz[r]});
                h128_of[r] = it == h128_idx.end() ? 0 : it->second;
            }
        co_await h128;
    }
    // starts of every big step-128 zone and their proxy test, one batch; then the proxy flood of each
    // zone's first passable start, one batch
    std::vector<std::vector<XY>> zone_starts(h128.jobs.size());
    PointsReq pxq{&S, seed, DSX_PRED_PROXY, pcut, {}, {}, {}, {}, {}};
    std::vector<size_t> px_off(h128.jobs.size() + 1, 0);
    std::optional<Timer> tz(16);
    for (size_t z = 0; z < h128.jobs.size(); z++) {
        const FloodOut &hzn = h128.out[z];
        const double area = hzn.a.cells * 128.0 * 128 * 16 / 1e6;
        px_off[z] = pxq.xs.size();
        if (area >= P.hot_min) {
            const int x0 = hzn.qx, z0 = hzn.qz, half = hzn.half;
            hzn.each_counted([&](int i, int j) {
                if (((i - half) & 3) == 0 && ((j - half) & 3) == 0) {
                    zone_starts[z].push_back({x0 + (i - half) * 128, z0 + (j - half) * 128});
                    pxq.xs.push_back(x0 + (i - half) * 128);
                    pxq.zs.push_back(z0 + (j - half) * 128);
                }
            });
        }
    }
    px_off[h128.jobs.size()] = pxq.xs.size();
    tz.reset();
    co_await pxq;
    // Proxy floods ahead of the loop, in rounds: each round floods, per zone, the first passable start
    // that no flood so far covers. A start the loop will flood is one that no earlier flood covers, so
    // after the rounds the loop almost never has to wait (the covered set it sees is a subset).
    std::unordered_map<XY, const FloodOut *, XYHash> prox_idx;
    std::vector<std::unique_ptr<FloodReq>> prox_rounds;
    KeySet spec_cov;
    // Each zone's primary sequence is what the loop below floods: its first passable start that no
    // primary flood covers, then the next, ... (the covered set only grows, so each zone keeps a cursor).
    // A round requests every zone's next primary start; a zone still open after r rounds also requests
    // up to 2^r - 1 further open starts (at most 7, DSX_PROXY_EXTRA), each 1024 quarts (DSX_PROXY_SEP)
    // from the others, as extras. An extra's result
    // only fills prox_idx: when the primary sequence reaches a start whose flood is already there, it is
    // taken without a round, so zones with many components need fewer rounds.
    std::vector<size_t> cursor(zone_starts.size(), 0);
    std::vector<int> zrounds(zone_starts.size(), 0);
    std::vector<uint8_t> is_primary;   // per request of the current round
    std::unordered_set<XY, XYHash> req_set;   // the starts the current round requests
    static const int extra_cap = int(env_u64("DSX_PROXY_EXTRA", 7)), extra_sep = int(env_u64("DSX_PROXY_SEP", 1024));
    for (;;) {
        std::optional<Timer> tr(17);
        auto req = std::make_unique<FloodReq>(FloodReq{&S, {}, true, {}, {}});
        is_primary.clear();
        req_set.clear();
        auto requested = [&](const XY &st) { return req_set.count(st) != 0; };
        for (size_t z = 0; z < zone_starts.size(); z++) {
            const size_t nz = zone_starts[z].size();
            size_t &k = cursor[z];
            auto open = [&](size_t j) {
                const XY st = zone_starts[z][j];
                return pxq.v8[px_off[z] + j] == 2 && !spec_cov.has(st.first, st.second);
            };
            // advance the primary sequence through starts already flooded
            for (; k < nz; k++) {
                if (!open(k)) continue;
                auto it = prox_idx.find(zone_starts[z][k]);
                if (it == prox_idx.end()) break;
                spec_cov.add(*it->second, 5, 1);
            }
            if (k >= nz) continue;
            const XY st = zone_starts[z][k];
            std::vector<XY> picks{st};
            if (!requested(st)) {
                req->jobs.push_back(fjob(seed, DSX_PRED_PROXY, pcut, st.first, st.second, 32, 16000 / 32, 1));
                req_set.insert(st);
                is_primary.push_back(1);
            } else {
                for (size_t j = 0; j < req->jobs.size(); j++)
                    if (req->jobs[j].qx == st.first && req->jobs[j].qz == st.second) is_primary[j] = 1;
            }
            const int extra = std::min((1 << std::min(zrounds[z], 16)) - 1, extra_cap);
            for (size_t j = k + 1; j < nz && int(picks.size()) <= extra; j++) {
                const XY sj = zone_starts[z][j];
                if (!open(j) || prox_idx.count(sj) || requested(sj)) continue;
                bool near = false;
                for (const XY &p : picks)
                    near |= std::max(std::abs(p.first - sj.first), std::abs(p.second - sj.second)) < extra_sep;
                if (near) continue;
                picks.push_back(sj);
                req->jobs.push_back(fjob(seed, DSX_PRED_PROXY, pcut, sj.first, sj.second, 32, 16000 / 32, 1));
                req_set.insert(sj);
                is_primary.push_back(0);
            }
            zrounds[z]++;
        }
        if (req->jobs.empty()) break;
        tr.reset();
        co_await *req;
        tr.emplace(17);
        for (size_t k = 0; k < req->jobs.size(); k++) {
            prox_idx[{req->jobs[k].qx, req->jobs[k].qz}] = &req->out[k];
            if (is_primary[k]) spec_cov.add(req->out[k], 5, 1);
        }
        prox_rounds.push_back(std::move(req));
    }

    std::optional<Timer> tl(18);
    for (size_t rank = 0; rank < ns; rank++) {
        const Site &st = sites[rank];
        std::vector<XY> starts;
        std::vector<int8_t> start_px;
        double hot_area = 0;
        if (P.hot_min > 0) {
            if (bt[rank] < 0.55 || hotcovered.has(hx[rank], hz[rank])) continue;
            if (P.hot_pre > 0) {
                const FloodOut &h2 = pre.out[pre_of[rank]];
                if (h2.a.cells * 256.0 * 256 * 16 / 1e6 < P.hot_pre * P.hot_min) {
                    hotcovered.add(h2, 7, 2);
                    O.nhot++;
                    continue;
                }
            }
            const size_t z = h128_of[rank];
            const FloodOut &hzn = h128.out[z];
            hot_area = hzn.a.cells * 128.0 * 128 * 16 / 1e6;
            hotcovered.add(hzn, 7, 1);
            O.nhot++;
            if (hot_area < P.hot_min) continue;
            O.nbig++;
            if (P.gpucheck > 0 && g_gpu_checked.fetch_add(1) < P.gpucheck) {
                // the zone's step-32 grid as dsearch_desert.cpp:367-371 forms it; every 7th cell (:447)
                const int x0 = int(std::floor((hzn.a.minx - 128) / 32.0)) * 32;
                const int z0 = int(std::floor((hzn.a.minz - 128) / 32.0)) * 32;
                const int gnx = (hzn.a.maxx + 128 - x0) / 32 + 1, gnz = (hzn.a.maxz + 128 - z0) / 32 + 1;
                PointsReq q{&S, seed, DSX_PRED_PROXY, pcut, {}, {}, {}, {}, {}};
                for (long c = 0; c < long(gnx) * gnz; c += 7) {
                    q.xs.push_back(x0 + int(c % gnx) * 32);
                    q.zs.push_back(z0 + int(c / gnx) * 32);
                }
                tl.reset();
                co_await q;
                tl.emplace(18);
                auto N = std::make_unique<cpu::Noise[]>(3);
                cpu::noise_init(seed, DSX_CH_T, N[0]);
                cpu::noise_init(seed, DSX_CH_C, N[1]);
                cpu::noise_init(seed, DSX_CH_E, N[2]);
                long diff = 0;
                for (size_t k = 0; k < q.xs.size(); k++)
                    diff += (cpu::proxy_pred(N[0], N[1], N[2], DSX_DPROB, pcut, q.xs[k], q.zs[k]) == 2) != (q.v8[k] == 2);
                std::fprintf(stderr, "gpucheck zone %dx%d: %ld/%zu cells differ\n", gnx, gnz, diff, q.xs.size());
            }
            starts = zone_starts[z];
            start_px.assign(pxq.v8.begin() + px_off[z], pxq.v8.begin() + px_off[z + 1]);
        } else {
            if (P.multi <= 1) {
                if (px_b[rank] >= pcut) starts.push_back({px_x[rank], px_z[rank]});
            } else {
                for (int dz = -P.multi; dz <= P.multi; dz++)
                    for (int dx = -P.multi; dx <= P.multi; dx++) starts.push_back({st.x + dx * 512, st.z + dz * 512});
            }
            if (starts.empty()) continue;
            PointsReq q{&S, seed, DSX_PRED_PROXY, pcut, {}, {}, {}, {}, {}};
            for (auto [bx, bz] : starts) {
                q.xs.push_back(bx);
                q.zs.push_back(bz);
            }
            tl.reset();
            co_await q;
            tl.emplace(18);
            start_px = q.v8;
        }
        // ---- proxy floods (dsearch_desert.cpp:384-400)
        for (size_t k = 0; k < starts.size(); k++) {
            const int bx = starts[k].first, bz = starts[k].second;
            if (start_px[k] != 2) continue;
            if (covered.has(bx, bz)) continue;
            const FloodOut *a;
            auto it = prox_idx.find(starts[k]);
            if (it != prox_idx.end()) {
                a = it->second;
            } else {
                FloodReq f{&S, {fjob(seed, DSX_PRED_PROXY, pcut, bx, bz, 32, 16000 / 32, 1)}, true, {}, {}};
                tl.reset();
                co_await f;
                tl.emplace(18);
                keep.push_back(std::make_unique<FloodOut>(std::move(f.out[0])));
                a = keep.back().get();
            }
            covered.add(*a, 5, 1);
            const std::pair<int, int> sig(a->a.minx, a->a.minz);
            if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
            seen.push_back(sig);
            Eval e{rank, bx, bz, a->a.cells * 32.0 * 32 * 16 / 1e6};
            e.hot = hot_area;
            e.px = a;
            ev.push_back(e);
        }
    }
    tl.reset();
    if (P.study && !ev.empty()) {   // before the sort: rows in the order the components were found
        FloodReq hz{&S, {}, true, {}, {}};
        for (const Eval &e : ev) hz.jobs.push_back(fjob(seed, DSX_PRED_HOT, 0.f, e.x, e.z, 64, 24000 / 64, 0));
        co_await hz;
        PointsReq qc{&S, seed, DSX_POINT_PROXY_C, 0.f, {}, {}, {}, {}, {}};
        std::vector<size_t> off;
        study_points(hz.out, qc, off);
        PointsReq qe{&S, seed, DSX_POINT_PROXY_E, 0.f, qc.xs, qc.zs, {}, {}, {}};
        co_await qc;
        co_await qe;
        study_rows(P, O, hz.out, qc.vd, qe.vd, off);
    }
    // ---- exact floods of the best proxies (dsearch_desert.cpp:494-533). The ev entries are
    // independent: ring searches of all of them in one batch, then their step-8 and step-4 floods.
    std::sort(ev.begin(), ev.end(), [](const Eval &a, const Eval &b) { return a.proxy > b.proxy; });
    size_t nx = 0;
    while (!P.study && nx < ev.size() && int(nx) < P.exact_n && ev[nx].proxy >= P.exact_min) nx++;
    // exact floods get evaluation hints (DSX_HINT_EXACT=0 turns them off): step 8 from the proxy component,
    // step 4 from the step-8 component
    static const bool hint_exact = env_u64("DSX_HINT_EXACT", 1) != 0;
    PointsReq ring{&S, seed, DSX_PRED_DESERT, 0.f, {}, {}, {}, {}, {}};
    for (size_t k2 = 0; k2 < nx; k2++)
        for (int r = 0; r <= 8; r++)
            for (int dz = -r; dz <= r; dz++)
                for (int dx = -r; dx <= r; dx++) {
                    if (std::max(std::abs(dx), std::abs(dz)) != r) continue;
                    ring.xs.push_back(ev[k2].x + dx * 16);
                    ring.zs.push_back(ev[k2].z + dz * 16);
                }
    co_await ring;
    const size_t per = ring.xs.size() / std::max<size_t>(nx, 1);
    FloodReq f8{&S, {}, hint_exact, {}, {}};   // counted cells: the step-4 floods' hints
    std::vector<int> f8_of(nx, -1);
    for (size_t k2 = 0; k2 < nx; k2++) {
        for (size_t k = k2 * per; k < (k2 + 1) * per; k++)
            if (ring.v8[k] == 2) {
                f8_of[k2] = int(f8.jobs.size());
                if (hint_exact && ev[k2].px)
                    proxy_hints(*ev[k2].px, uint32_t(f8.jobs.size()), ring.xs[k], ring.zs[k], 16000 / 8, f8.hints);
                f8.jobs.push_back(fjob(seed, DSX_PRED_DESERT_STATE, 0.f, ring.xs[k], ring.zs[k], 8, 16000 / 8, 1));
                break;
            }
    }
    co_await f8;
    FloodReq f4{&S, {}, false, {}, {}};
    std::vector<int> f4_of(nx, -1);
    for (size_t k2 = 0; k2 < nx; k2++) {
        if (f8_of[k2] < 0) continue;
        const FloodOut &a8 = f8.out[f8_of[k2]];
        if (a8.a.cells * 64.0 * 16 / 1e6 >= P.refine_mb) {
            f4_of[k2] = int(f4.jobs.size());
            if (hint_exact) refine_hints(a8, uint32_t(f4.jobs.size()), 16000 / 4, f4.hints);
            f4.jobs.push_back(fjob(seed, DSX_PRED_DESERT_STATE, 0.f, a8.qx, a8.qz, 4, 16000 / 4, 0));
        }
    }
    co_await f4;
    double seed_best = 0;
    for (size_t k2 = 0; k2 < nx; k2++) {
        Eval &e = ev[k2];
        if (f8_of[k2] < 0) {
            e.s8 = 0;
            continue;
        }
        const FloodOut &a8 = f8.out[f8_of[k2]];
        const int bx = a8.qx, bz = a8.qz;
        e.x = bx;
        e.z = bz;
        e.s8 = a8.a.cells * 64.0 * 16 / 1e6;
        e.edge = a8.a.edge != 0;
        if (f4_of[k2] >= 0) {
            const FloodOut &a4 = f4.out[f4_of[k2]];
            e.s4 = a4.a.cells * 16.0 * 16 / 1e6;
            e.edge = e.edge || a4.a.edge != 0;
        }
        const double v = e.s4 > 0 ? e.s4 : e.s8;
        seed_best = std::max(seed_best, v);
        if (e.s4 >= P.verify_mb) {
            char line[256];
            std::snprintf(line, sizeof line, "CANDIDATE step4=%.2fM seed=%lld x=%d z=%d -> verify with exact flood (Y=256)",
                          e.s4, (long long)seed, bx * 4 + 2, bz * 4 + 2);
            O.lines.push_back(line);
            O.cand_xy.push_back({bx * 4 + 2, bz * 4 + 2});
        }
        O.bests.push_back(v);
        O.best_xy.push_back({bx * 4 + 2, bz * 4 + 2});
        O.best_proxy.push_back(e.proxy);
        O.best_step.push_back(e.s4 > 0 ? 4 : 8);
    }
    O.best = seed_best;
    O.done = true;
}

// ------------------------------------------------------------------ seeds -> sites (S1-S2 on the GPU)
// dsearch_desert.cpp:227, 274-277, 301-308
void make_sites(const dsx_raw_cand *raw, size_t nraw, const dsx_search_params &P, SeedOut &O) {
    std::vector<Site> cand;
    std::unordered_map<uint64_t, uint32_t> bucket;
    for (size_t i = 0; i < nraw; i++) {
        const dsx_raw_cand &r = raw[i];
        const uint64_t key = (uint64_t(uint32_t((r.x + 8000000) / P.nms)) << 32) | uint32_t((r.z + 8000000) / P.nms);
        auto [it, fresh] = bucket.try_emplace(key, uint32_t(cand.size()));
        if (fresh) cand.push_back({r.s, r.t, r.c, r.e, r.x, r.z});
        else if (r.s > cand[it->second].s) cand[it->second] = {r.s, r.t, r.c, r.e, r.x, r.z};
    }
    O.ncand = cand.size();
    std::sort(cand.begin(), cand.end(), [](const Site &a, const Site &b) { return a.s > b.s; });
    // a candidate is dropped when a kept site lies within nms on both axes; kept sites are found by
    // nms-sized cell, so only the 3 x 3 cells around the candidate can hold one
    std::unordered_map<uint64_t, std::vector<uint32_t>> kept;
    auto cell = [&](int v) { return uint32_t((int64_t(v) + 8000000) / P.nms); };
    for (auto &c : cand) {
        bool near = false;
        const uint32_t cx = cell(c.x), cz = cell(c.z);
        for (uint32_t gz = cz - 1; gz != cz + 2 && !near; gz++)
            for (uint32_t gx = cx - 1; gx != cx + 2 && !near; gx++) {
                const auto it = kept.find((uint64_t(gx) << 32) | gz);
                if (it == kept.end()) continue;
                for (const uint32_t k : it->second) {
                    const Site &q = O.sites[k];
                    if (std::abs(q.x - c.x) < P.nms && std::abs(q.z - c.z) < P.nms) { near = true; break; }
                }
            }
        if (!near) {
            kept[(uint64_t(cx) << 32) | cz].push_back(uint32_t(O.sites.size()));
            O.sites.push_back(c);
        }
        if (int(O.sites.size()) >= P.max_sites) break;
    }
}

static void gscan_check(uint64_t seed, const dsx_search_params &P, float gmaxta, float gmaxtb, const dsx_raw_cand *raw,
                        size_t nraw) {
    const cpu::ScanResult R = cpu::scan_seed(seed, P.sigma, P.sigmac, float(P.ta), float(P.t), float(P.wc), float(P.we),
                                             P.nms);
    std::vector<cpu::Cand> gc;
    std::unordered_map<uint64_t, uint32_t> gb;
    for (size_t i = 0; i < nraw; i++) {
        const dsx_raw_cand &r = raw[i];
        auto [it, fresh] = gb.try_emplace(cpu::bucket_key(r.x, r.z, P.nms), uint32_t(gc.size()));
        if (fresh) gc.push_back({r.s, r.t, r.c, r.e, r.x, r.z});
        else if (r.s > gc[it->second].s) gc[it->second] = {r.s, r.t, r.c, r.e, r.x, r.z};
    }
    long miss = 0;
    for (const cpu::Cand &c : R.cand) {
        auto it = gb.find(cpu::bucket_key(c.x, c.z, P.nms));
        if (it == gb.end() || gc[it->second].s != c.s) miss++;
    }
    std::fprintf(stderr,
                 "gscancheck seed %lld: maxta %.6f/%.6f maxtb %.6f/%.6f, buckets cpu %zu gpu %zu, mismatched %ld, raw %zu\n",
                 (long long)seed, R.maxta, gmaxta, R.maxtb, gmaxtb, R.cand.size(), gc.size(), miss, nraw);
}

// ------------------------------------------------------------------ groups of seeds on worker contexts
Group::Group() = default;
Group::~Group() = default;

// Tables, scan and sites of a group's seeds (g.outs).
std::atomic<uint64_t> g_cand_cap{1u << 12};
int group_scan(dsx_ctx *w, const dsx_search_params &P, int mode, Group &g) {
    const dsx_scan_params sp{float(P.ta), float(P.t), float(P.wc), float(P.we), P.sigma, P.sigmac};
    const size_t nb = g.seeds.size();
    auto tscan = std::chrono::steady_clock::now();
    std::vector<float> mta(nb), mtb(nb);
    std::vector<uint64_t> ncand(nb);
    // room per seed: the most candidates any seed of the process has had, up to 2^15 (a seed with more
    // makes dsx_scan report the capacity, and the group is scanned again with room for it)
    uint64_t cap = g_cand_cap.load(std::memory_order_relaxed);
    thread_local std::vector<dsx_raw_cand> raw;   // reused: dsx_scan writes the ncand[k] entries read below
    if (raw.size() < nb * cap) raw.resize(nb * cap);
    int rc = dsx_scan(w, g.seeds.data(), nb, &sp, mta.data(), mtb.data(), raw.data(), cap, ncand.data());
    if (rc == DSX_E_CAPACITY) {
        cap = std::max(cap, *std::max_element(ncand.begin(), ncand.end()));
        const uint64_t keep = std::min<uint64_t>(cap, 1u << 15);
        uint64_t seen = g_cand_cap.load(std::memory_order_relaxed);
        while (seen < keep && !g_cand_cap.compare_exchange_weak(seen, keep, std::memory_order_relaxed)) {
        }
        if (raw.size() < nb * cap) raw.resize(nb * cap);
        rc = dsx_scan(w, g.seeds.data(), nb, &sp, mta.data(), mtb.data(), raw.data(), cap, ncand.data());
    }
    if (rc) return rc;
    for (size_t k = 0; k < nb; k++) {
        auto O = std::make_unique<SeedOut>();
        O->seed = g.seeds[k];
        dsx::temperature_phase(O->seed, O->fy);   // fyA, fyB (dsearch_desert.cpp:80-94)
        O->maxta = mta[k];
        O->maxtb = mtb[k];
        if (P.gscancheck > 0 && g_gscan_checked.fetch_add(1) < P.gscancheck)
            gscan_check(O->seed, P, mta[k], mtb[k], raw.data() + k * cap, ncand[k]);
        if (!(mode == DSX_SEARCH_ENUM && O->maxta + O->maxtb < P.amp)) {
            O->scanned = true;
            Timer tm(8);
            make_sites(raw.data() + k * cap, ncand[k], P, *O);
        }
        g.outs.push_back(std::move(O));
    }
    prof_add(1, std::chrono::duration<double>(std::chrono::steady_clock::now() - tscan).count(), 0);
    return DSX_OK;
}

// Hot zones, proxy floods, exact floods of every scanned seed as coroutines in S: each starts and runs
// to its first request.
int group_tasks(Sched &S, const dsx_search_params &P, Group &g, std::vector<Task> &tasks) {
    if (const int rc = dsx_set_precision(S.ctx, P.prepass ? 1 : 0)) return rc;
    for (auto &O : g.outs)
        if (O->scanned) tasks.push_back(seed_task(S, P, *O));
    return DSX_OK;
}

// After every seed of the group finished: the in-process verification (mcverify=1), the exact step-1
// area of every CANDIDATE in one batch.
int group_finish(dsx_ctx *w, const dsx_search_params &P, Group &g) {
    if (!P.mcverify) return DSX_OK;
    std::vector<dsx_area_query> q;
    for (auto &O : g.outs)
        for (auto [x, z] : O->cand_xy) q.push_back(dsx_area_query{O->seed, x, z, 20000, 0});
    if (q.empty()) return DSX_OK;
    std::vector<dsx_area> a(q.size());
    if (const int rc = dsx_area_queries(w, q.data(), q.size(), a.data())) return rc;
    size_t k = 0;
    for (auto &O : g.outs)
        for (size_t c = 0; c < O->cand_xy.size(); c++) O->cand_area.push_back(a[k++]);
    return DSX_OK;
}

int run_group(dsx_ctx *w, const dsx_search_params &P, int mode, Group &g) {
    Sched S{w, {}, {}, DSX_OK, {}};
    std::vector<Task> tasks;
    if (const int rc = group_scan(w, P, mode, g)) return rc;
    if (const int rc = group_tasks(S, P, g, tasks)) return rc;
    auto tsites = std::chrono::steady_clock::now();
    while (S.step()) {
    }
    for (auto &t : tasks) t.h.destroy();
    prof_add(2, std::chrono::duration<double>(std::chrono::steady_clock::now() - tsites).count(), 0);
    if (S.err) return S.err;
    return group_finish(w, P, g);
}

void format_group(const Group &g, const dsx_search_params &P, GroupText &t, double *best) {
    char line[512];
    for (auto &O : g.outs) {
        for (size_t c = 0; c < O->cand_area.size(); c++) {   // "<CANDIDATE line> => <MC-Measure result>"
            std::snprintf(line, sizeof line, " => 1:4 Size: %llu | Blocks^2 Size: %llu\n",
                          (unsigned long long)O->cand_area[c].cells, (unsigned long long)(16 * O->cand_area[c].cells));
            t.verified.push_back(O->lines[c] + line);
        }
        if (!O->scanned) {
            std::snprintf(line, sizeof line, "%lld,%.4f,%.4f,%.4f,%.4f,0,0,0\n", (long long)O->seed, O->fy[0], O->fy[1],
                          O->maxta, O->maxtb);
            t.seeds += line;
            continue;
        }
        t.scanned++;
        t.sites += O->sites.size();
        t.hot_zones += O->nhot;
        t.big_zones += O->nbig;
        for (auto &l : O->lines) t.log += l + "\n";
        for (size_t k = 0; k < O->bests.size(); k++)
            if (best && O->bests[k] > *best) {
                *best = O->bests[k];
                std::snprintf(line, sizeof line, "best %.2fM (step %d) seed=%lld x=%d z=%d proxy=%.1fM\n", O->bests[k],
                              O->best_step[k], (long long)O->seed, O->best_xy[k].first, O->best_xy[k].second,
                              O->best_proxy[k]);
                t.log += line;
            }
        for (auto &e : O->ev) {
            if (e.proxy < P.csv_min && e.s8 <= 0 && e.s4 <= 0) continue;
            const Site &st = O->sites[e.rank];
            std::snprintf(line, sizeof line, "%lld,%.4f,%.4f,%.4f,%.4f,%zu,%d,%d,%.4f,%.4f,%.4f,%.4f,%.3f,%.3f,%.3f,%d,%.1f\n",
                          (long long)O->seed, O->fy[0], O->fy[1], O->maxta, O->maxtb, e.rank, e.x * 4 + 2, e.z * 4 + 2, st.s,
                          st.t, st.c, st.e, e.proxy, e.s8, e.s4, int(e.edge), e.hot);
            t.csv += line;
        }
        std::snprintf(line, sizeof line, "%lld,%.4f,%.4f,%.4f,%.4f,%zu,%zu,%.3f\n", (long long)O->seed, O->fy[0], O->fy[1],
                      O->maxta, O->maxtb, O->ncand, O->sites.size(), O->best);
        t.seeds += line;
        for (auto &z : O->zones) t.zones += z;
    }
}

// Tuning knobs from the environment (DSX_CHUNK: seed indices gated per chunk, DSX_GROUP: seeds per
// worker group).
uint64_t env_u64(const char *name, uint64_t def) {
    const char *v = std::getenv(name);
    return v && *v ? std::strtoull(v, nullptr, 10) : def;
}

// Worker contexts of a caller's context, created on first use and kept until dsx_destroy.
std::mutex g_workers_mu;
std::unordered_map<dsx_ctx *, std::vector<dsx_ctx *>> g_workers;

int worker_contexts(dsx_ctx *ctx, int n, std::vector<dsx_ctx *> &out) {
    std::lock_guard<std::mutex> l(g_workers_mu);
    auto &v = g_workers[ctx];
    while (int(v.size()) < n) {
        dsx_ctx *w = nullptr;
        const int rc = dsx_create(dsx_ctx_device(ctx), &w);
        if (rc) return rc;
        v.push_back(w);
    }
    out.assign(v.begin(), v.begin() + n);
    return DSX_OK;
}

// Runs groups on worker threads; the caller submits groups and collects them in submission order.
// With scan workers, a group's tables, scan and sites run on a scan worker's context, then its floods
// on a flood worker's, so the flood workers' device calls are floods and points only while the next
// groups are scanned.
class Pool {
  public:
    Pool(const std::vector<dsx_ctx *> &flood_ws, const std::vector<dsx_ctx *> &scan_ws, const dsx_search_params &P,
         int mode)
        : P_(P), mode_(mode), nscan_(scan_ws.size()) {
        for (dsx_ctx *w : flood_ws) th_.emplace_back([this, w] { loop(w); });
        for (dsx_ctx *w : scan_ws) th_.emplace_back([this, w] { scan_loop(w); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> l(mu_);
            closing_ = true;
        }
        cv_work_.notify_all();
        for (auto &t : th_) t.join();
    }
    void submit(Group *g) {
        {
            std::lock_guard<std::mutex> l(mu_);
            (nscan_ ? qs_ : q_).push_back(g);
        }
        cv_work_.notify_all();
    }
    void wait(Group *g) {
        std::unique_lock<std::mutex> l(mu_);
        cv_done_.wait(l, [g] { return g->done; });
    }
    bool ready(Group *g) {
        std::lock_guard<std::mutex> l(mu_);
        return g->done;
    }
    void stop() {   // groups not yet started finish at once with an error
        std::lock_guard<std::mutex> l(mu_);
        stop_ = true;
    }

  private:
    // no group can reach the flood queue any more (caller holds mu_)
    bool drained() const { return closing_ && qs_.empty() && scanning_ == 0; }

    void scan_loop(dsx_ctx *w) {
        for (;;) {
            Group *g;
            bool skip;
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_work_.wait(l, [this] { return closing_ || !qs_.empty(); });
                if (qs_.empty()) return;
                g = qs_.front();
                qs_.pop_front();
                skip = stop_;
                scanning_++;
            }
            const int rc = skip ? DSX_E_ARG : group_scan(w, P_, mode_, *g);
            const std::string err = skip ? "search stopped after an earlier error" : (rc ? dsx_last_error() : "");
            {
                std::lock_guard<std::mutex> l(mu_);
                scanning_--;
                if (!rc) {
                    g->scanned = true;
                    q_.push_back(g);
                }
            }
            if (rc) finish(g, rc, err);
            cv_work_.notify_all();
        }
    }

    void finish(Group *g, int rc, const std::string &err) {
        {
            std::lock_guard<std::mutex> l(mu_);
            g->rc = rc;
            g->err = err;
            g->done = true;
        }
        cv_done_.notify_all();
    }
    // Continuous batching (DSX_ACTIVE > 1): the worker keeps the seeds of several groups in one
    // scheduler, so every flood and point call carries the requests of all seeds in flight, and a new
    // group is admitted whenever fewer than DSX_ACTIVE seeds are in flight. Results do not depend on the
    // mix: every request is a pure function of its seed and arguments. The default, 1, runs one group
    // at a time: with 16 workers the device is the bottleneck, and larger batches only made each call
    // longer (measured 9.5-9.9k against 9.8-10.0k gated seeds/s).
    void loop(dsx_ctx *w) {
        static const size_t active = size_t(env_u64("DSX_ACTIVE", 1));
        Sched S{w, {}, {}, DSX_OK, {}};
        struct Active {
            Group *g;
            std::vector<Task> tasks;
        };
        std::deque<Active> act;
        size_t nact = 0;
        for (;;) {
            while (nact < active) {   // admit groups (wait for one only when nothing is in flight)
                Group *g;
                bool skip;
                {
                    std::unique_lock<std::mutex> l(mu_);
                    if (act.empty()) cv_work_.wait(l, [this] { return !q_.empty() || drained(); });
                    if (q_.empty()) break;
                    g = q_.front();
                    q_.pop_front();
                    skip = stop_;
                }
                if (skip) {
                    finish(g, DSX_E_ARG, "search stopped after an earlier error");
                    continue;
                }
                Active a{g, {}};
                int rc = g->scanned ? DSX_OK : group_scan(w, P_, mode_, *g);
                if (!rc) rc = group_tasks(S, P_, *g, a.tasks);
                if (rc) {   // the scan failed: no coroutine started
                    finish(g, rc, dsx_last_error());
                    continue;
                }
                nact += a.tasks.size();
                act.push_back(std::move(a));
            }
            if (act.empty()) {
                std::lock_guard<std::mutex> l(mu_);
                if (q_.empty() && drained()) return;   // closing, nothing left
                continue;
            }
            {
                const auto ts = std::chrono::steady_clock::now();   // per-seed site stages (hot, proxy, exact)
                S.step();
                prof_add(2, std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count(), 0);
            }
            if (S.err) {   // a device call failed: every group in flight fails with it
                const std::string err = dsx_last_error();
                for (auto &a : act) {
                    for (auto &t : a.tasks) t.h.destroy();
                    finish(a.g, S.err, err);
                }
                act.clear();
                nact = 0;
                S.points.clear();
                S.floods.clear();
                S.err = DSX_OK;
                continue;
            }
            for (auto it = act.begin(); it != act.end();) {   // retire the groups whose seeds all finished
                bool all = true;
                for (auto &t : it->tasks) all = all && t.h.done();
                if (!all) {
                    ++it;
                    continue;
                }
                for (auto &t : it->tasks) t.h.destroy();
                nact -= it->tasks.size();
                const int rc = group_finish(w, P_, *it->g);
                finish(it->g, rc, rc ? dsx_last_error() : "");
                it = act.erase(it);
            }
        }
    }
    const dsx_search_params P_;
    const int mode_;
    const size_t nscan_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Group *> q_, qs_;   // groups ready for floods; groups to scan
    size_t scanning_ = 0;          // groups a scan worker holds
    bool closing_ = false, stop_ = false;
    std::vector<std::thread> th_;
};

}  // namespace dsxh

using namespace dsxh;

extern "C" {

void dsx_search_defaults(dsx_search_params *p) {
    dsx_search_params d{};
    d.max_phase = 0.10; d.amp = 0.70; d.amp_margin = 0.01; d.sigma = 2.0; d.sigmac = 2.0;
    d.wc = 0.25; d.we = 0.41; d.t = 0.55; d.ta = 0.25; d.max_sites = 300; d.pcut = 0.35; d.hot_min = 180;
    d.multi = 1; d.exact_n = 2; d.exact_min = 75; d.hot_pre = 0.8; d.csv_min = 50; d.nms = 5000;
    d.refine_mb = 70; d.verify_mb = 80; d.study = 0; d.gpu = 0; d.gpucheck = 0; d.gscan = 0; d.gscancheck = 0;
    d.gate = 0; d.gate_phase = 0.06; d.gate_amp = 0.70; d.gate_pthr = 0.645; d.from_stdin = 0; d.workers = 0;
    d.prepass = 0; d.mcverify = 0;
    *p = d;
}

int dsx_interrupt(int on) {
    g_interrupt.store(on ? 1 : 0, std::memory_order_relaxed);
    return DSX_OK;
}

int dsx_search(dsx_ctx *ctx, const dsx_search_params *pp, int mode, uint64_t start, uint64_t count, const uint64_t *seeds,
               uint64_t nseeds, const char *prefix, dsx_search_counts *counts) {
    if (!ctx || !pp || !prefix || !counts || mode < 0 || mode > 2 || (mode == DSX_SEARCH_LIST && nseeds && !seeds))
        return DSX_E_ARG;
    if (g_interrupt.load(std::memory_order_relaxed)) return DSX_E_INTERRUPTED;
    double prof0[20];
    {
        std::lock_guard<std::mutex> l(g_prof_mu);
        std::copy(g_prof.t, g_prof.t + 20, prof0);
    }
    const dsx_search_params P = *pp;
    if (P.nms <= 0 || P.max_sites < 0 || P.workers < 0 || P.workers > 64) return DSX_E_ARG;
    // workers = 0: DSX_WORKERS, else 16 (one per host core of a 16-core share)
    const int nw = P.workers ? P.workers : int(std::clamp<uint64_t>(env_u64("DSX_WORKERS", 16), 1, 64));
    std::vector<dsx_ctx *> wctx;
    const int ns = int(std::min<uint64_t>(env_u64("DSX_SCAN_WORKERS", 0), 8));
    int rc = worker_contexts(ctx, nw + ns, wctx);
    if (rc) return rc;
    const std::vector<dsx_ctx *> flood_ws(wctx.begin(), wctx.begin() + nw), scan_ws(wctx.begin() + nw, wctx.end());
    const std::string pre(prefix);
    FILE *out = std::fopen((pre + ".csv").c_str(), "a");
    FILE *sout = std::fopen((pre + ".seeds.csv").c_str(), "a");
    if (!out || !sout) {
        if (out) std::fclose(out);
        if (sout) std::fclose(sout);
        dsx::g_last_error = "cannot open " + pre + ".csv / .seeds.csv";
        return DSX_E_ARG;
    }
    FILE *vout = P.mcverify ? std::fopen((pre + ".verified").c_str(), "a") : nullptr;
    FILE *zout = P.study ? std::fopen((pre + ".zones.csv").c_str(), "a") : nullptr;
    if (P.study && !zout) {
        std::fclose(out);
        std::fclose(sout);
        if (vout) std::fclose(vout);
        dsx::g_last_error = "cannot open " + pre + ".zones.csv";
        return DSX_E_ARG;
    }
    const size_t slash = pre.find_last_of('/');
    const std::string run = slash == std::string::npos ? pre : pre.substr(slash + 1);
    static double global_best = 0;
    std::deque<std::unique_ptr<Group>> pending;   // submitted, in seed order
    std::string first_err;
    {
        Pool pool(flood_ws, scan_ws, P, mode);
        // write the finished groups at the front, waiting for them while more than `keep` are pending
        auto drain = [&](size_t keep) {
            while (!pending.empty()) {
                Group *g = pending.front().get();
                if (pending.size() > keep) pool.wait(g);
                else if (!pool.ready(g)) break;
                if (g->rc && rc == DSX_OK) {
                    rc = g->rc;
                    first_err = g->err;
                    pool.stop();
                }
                if (rc == DSX_OK) {
                    GroupText t;
                    format_group(*g, P, t, &global_best);
                    std::fputs(t.csv.c_str(), out);
                    std::fputs(t.seeds.c_str(), sout);
                    if (zout) std::fputs(t.zones.c_str(), zout);
                    std::fputs(t.log.c_str(), stderr);
                    if (vout && !t.verified.empty()) {
                        char ts[32];
                        const std::time_t now = std::time(nullptr);
                        std::strftime(ts, sizeof ts, "%Y-%m-%dT%H:%M:%SZ", std::gmtime(&now));
                        for (const std::string &v : t.verified) std::fprintf(vout, "%s %s %s", ts, run.c_str(), v.c_str());
                        std::fflush(vout);
                    }
                    counts->scanned += t.scanned;
                    counts->sites += t.sites;
                    counts->hot_zones += t.hot_zones;
                    counts->big_zones += t.big_zones;
                    if (mode != DSX_SEARCH_ENUM) {   // gated seeds count as they are written
                        counts->done += g->seeds.size();
                        counts->passed += g->seeds.size();
                    }
                }
                pending.pop_front();
            }
        };
        // seed source in chunks; the gate of the next chunk overlaps the workers' groups of this one
        const uint64_t CHUNK = env_u64("DSX_CHUNK", 1ull << 22);
        const size_t group = size_t(std::clamp<uint64_t>(env_u64("DSX_GROUP", 64), 1, 4096));
        const uint64_t total = mode == DSX_SEARCH_LIST ? nseeds : count;
        std::vector<uint64_t> batch;
        bool stopped = false;
        for (uint64_t pos = 0; pos < total && rc == DSX_OK && !stopped;) {
            const uint64_t m = std::min(CHUNK, total - pos);
            batch.clear();
            auto tgate = std::chrono::steady_clock::now();
            if (g_interrupt.load(std::memory_order_relaxed)) {
                stopped = true;
                break;
            }
            if (mode == DSX_SEARCH_LIST) {
                batch.assign(seeds + pos, seeds + pos + m);
            } else if (mode == DSX_SEARCH_GATE) {
                dsx_gate_params gp{float(P.gate_phase), float(P.gate_amp), float(P.gate_pthr), 0.01f, 1.0f, 1};
                uint64_t n = 0;
                dsx_gate_counts gc;
                batch.resize(1 << 16);
                int r = dsx_gate(ctx, start + pos, m, &gp, batch.data(), nullptr, nullptr, batch.size(), &n, &gc);
                if (r == DSX_E_CAPACITY) {
                    batch.resize(n);
                    r = dsx_gate(ctx, start + pos, m, &gp, batch.data(), nullptr, nullptr, batch.size(), &n, &gc);
                }
                if (r) { rc = r; first_err = dsx_last_error(); break; }
                batch.resize(n);
            } else {
                uint64_t n = 0, nph = 0;
                batch.resize(1 << 18);
                int r = dsx_gate_cpu(ctx, start + pos, m, P.max_phase, P.amp, P.amp_margin, batch.data(), nullptr, nullptr,
                                     batch.size(), &n, &nph);
                if (r == DSX_E_CAPACITY) {
                    batch.resize(n);
                    r = dsx_gate_cpu(ctx, start + pos, m, P.max_phase, P.amp, P.amp_margin, batch.data(), nullptr, nullptr,
                                     batch.size(), &n, &nph);
                }
                if (r) { rc = r; first_err = dsx_last_error(); break; }
                batch.resize(n);
                counts->done += m;
                counts->passed += nph;
            }
            pos += m;
            prof_add(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - tgate).count(), 0);
            for (size_t b0 = 0; b0 < batch.size(); b0 += group) {
                // stop between groups: those in flight finish and are written (the CPU-gate mode finishes its
                // chunk, whose counts are already taken)
                if (mode != DSX_SEARCH_ENUM && g_interrupt.load(std::memory_order_relaxed)) {
                    stopped = true;
                    break;
                }
                auto g = std::make_unique<Group>();
                g->seeds.assign(batch.begin() + b0, batch.begin() + std::min(batch.size(), b0 + group));
                pool.submit(g.get());
                pending.push_back(std::move(g));
            }
            drain(size_t(8 * nw));   // bounded backlog; rows of finished groups go out early
        }
        drain(0);
        if (rc == DSX_OK && stopped) {
            rc = DSX_E_INTERRUPTED;
            first_err = "interrupted";
        }
    }
    std::fclose(out);
    std::fclose(sout);
    if (vout) std::fclose(vout);
    if (zout) std::fclose(zout);
    {
        std::lock_guard<std::mutex> l(g_prof_mu);
        const double *t = g_prof.t;
        // Timer categories: 1 tables + scan stage, 8 make_sites, 2 per-seed site stages, 3 point batches,
        // 4 flood batches, 16 hot-zone starts, 17 proxy-flood rounds
        counts->sec_tables += (t[1] - prof0[1]) - (t[8] - prof0[8]);
        counts->sec_scan += t[8] - prof0[8];
        counts->sec_sites += t[2] - prof0[2];
        counts->sec_proxy += t[17] - prof0[17];
        counts->sec_hot += t[16] - prof0[16];
        counts->sec_gpu += (t[3] - prof0[3]) + (t[4] - prof0[4]);
    }
    if (rc) dsx::g_last_error = first_err;
    return rc;
}

}  // extern "C"

void dsx_ctx_free_workers(dsx_ctx *c) {
    std::vector<dsx_ctx *> v;
    {
        std::lock_guard<std::mutex> l(g_workers_mu);
        auto it = g_workers.find(c);
        if (it == g_workers.end()) return;
        v.swap(it->second);
        g_workers.erase(it);
    }
    for (dsx_ctx *w : v) dsx_destroy(w);
}
