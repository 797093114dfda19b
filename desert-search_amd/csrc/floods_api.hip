// C ABI for the climate sampler and the flood engine (include/dsx.h: dsx_noise, dsx_climate,
// dsx_predicates, dsx_flood). Replaces the CPU floods of dsearch_desert.cpp:326-533 and tools/dflood.cpp.
#include "climate_api.h"
#include "dsx_rng.h"
#include "ctx.h"
#include "floods.h"
#include "../host/pipeline.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

// Surface-biome grid (surface.hpp:12-55): the depth-0 leaves of the parameter tree are boxes that
// tile (T, H, C, E, W) space; index them on the grid of all box boundaries. A cell covered by one
// biome maps to it; cells covered by several biomes or none map to -1 (the tree decides).
struct HostGrid {
    std::vector<int32_t> cuts[5];
    std::vector<int16_t> cell;
    int dim[5] = {};
    HostGrid() {
        static const int AX[5] = {0, 1, 2, 3, 5};
        struct Box { int32_t lo[5], hi[5]; int biome; };
        std::vector<Box> boxes;
        for (int n = 0; n < DSX_NTREE; n++) {
            const uint64_t e = DSX_TREE[n];
            if ((e >> 56) != 255) continue;
            const int32_t *d = DSX_PARAM[(e >> 32) & 255];
            if (d[0] != 0 || d[1] != 0) continue;
            Box b;
            b.biome = int((e >> 48) & 255);
            for (int a = 0; a < 5; a++) {
                const int32_t *r = DSX_PARAM[(e >> (8 * AX[a])) & 255];
                b.lo[a] = r[0];
                b.hi[a] = r[1];
            }
            boxes.push_back(b);
        }
        for (int a = 0; a < 5; a++) {
            for (const Box &b : boxes) { cuts[a].push_back(b.lo[a]); cuts[a].push_back(b.hi[a] + 1); }
            std::sort(cuts[a].begin(), cuts[a].end());
            cuts[a].erase(std::unique(cuts[a].begin(), cuts[a].end()), cuts[a].end());
            dim[a] = int(cuts[a].size()) - 1;
        }
        size_t total = 1;
        for (int a = 0; a < 5; a++) total *= size_t(dim[a]);
        cell.assign(total, -2);
        for (const Box &b : boxes) {
            int lo[5], hi[5];
            for (int a = 0; a < 5; a++) {
                lo[a] = int(std::lower_bound(cuts[a].begin(), cuts[a].end(), b.lo[a]) - cuts[a].begin());
                hi[a] = int(std::lower_bound(cuts[a].begin(), cuts[a].end(), b.hi[a] + 1) - cuts[a].begin());
            }
            for (int i0 = lo[0]; i0 < hi[0]; i0++)
                for (int i1 = lo[1]; i1 < hi[1]; i1++)
                    for (int i2 = lo[2]; i2 < hi[2]; i2++)
                        for (int i3 = lo[3]; i3 < hi[3]; i3++)
                            for (int i4 = lo[4]; i4 < hi[4]; i4++) {
                                const size_t k = ((((size_t(i0) * dim[1] + i1) * dim[2] + i2) * dim[3] + i3) * dim[4] + i4);
                                int16_t &v = cell[k];
                                v = (v == -2) ? int16_t(b.biome) : (v == b.biome ? v : int16_t(-1));
                            }
        }
        for (int16_t &v : cell)
            if (v == -2) v = -1;
    }
};

}  // namespace

// Device-resident climate state of a context (allocated on first use).
// Control blocks in rotation: a block is zeroed two batches before its use, so the zeroing never
// shares a batch boundary with the block's own batch.
constexpr uint32_t CTL_BLOCKS = 4;

struct ClimState {
    dsx::SurfGrid grid{};
    int32_t *d_cuts = nullptr;
    int16_t *d_cell = nullptr;
    dsx::ClimSeed *d_clim = nullptr;
    uint32_t clim_cap = 0;
    uint64_t *h_useeds = nullptr;   // pinned: seeds to initialise, then their slots (u32), read by k_climinit
    std::unordered_map<uint64_t, uint32_t> cslot;   // seed -> climate slot (initialised on the device)
    std::vector<uint64_t> miss;
    // flood buffers
    dsx::FloodBufs fb{};
    dsx::FloodRes *d_res = nullptr;
    uint32_t job_cap = 0;
    // batch I/O staging
    void *d_io = nullptr;
    void *h_io = nullptr;
    size_t io_cap = 0;
    // counted-cell tiles of the last dsx_flood call
    dsx::TileMask *d_masks = nullptr;
    uint32_t mask_cap = 0;
    std::vector<dsx::TileMask> masks;   // masks beyond the pinned buffer's capacity
    uint32_t masks_n = 0, masks_h = 0;  // masks of the last call; how many of them are in h_masks
    std::vector<hipEvent_t> events;   // event pair around the tiled engine's kernel (profiling on)
    int prepass = 0;                  // dsx_set_precision: 1 = FP32 pre-pass in floods and predicates
    uint32_t *d_ctl = nullptr;        // CTL_BLOCKS blocks of control words; batch i uses block i % CTL_BLOCKS
    uint32_t ctl_seq = 0;
    // flood batch I/O: jobs and job lists go up in one copy from pinned staging (device layout: n jobs,
    // n small-window indices, n tiled-engine indices); results come down through pinned memory the
    // publish kernel writes
    uint8_t *d_up = nullptr;
    uint8_t *h_up = nullptr;
    size_t h_up_cap = 0;
    dsx::FloodRes *h_res = nullptr;
    uint32_t *h_ctl = nullptr;
    dsx::TileMask *h_masks = nullptr;
    uint32_t h_res_cap = 0, h_mask_cap = 0, mask_want = 0;
};

static ClimState *clim_state(dsx_ctx *c);

static int clim_alloc(dsx_ctx *c, ClimState *&S) {
    S = clim_state(c);
    if (S->d_cell) return DSX_OK;
    static const HostGrid G;
    size_t ncut = 0;
    for (int a = 0; a < 5; a++) ncut += G.cuts[a].size();
    HIPCK(hipMalloc(&S->d_cuts, ncut * 4));
    HIPCK(hipMalloc(&S->d_cell, G.cell.size() * 2));
    size_t off = 0;
    for (int a = 0; a < 5; a++) {
        HIPCK(hipMemcpyAsync(S->d_cuts + off, G.cuts[a].data(), G.cuts[a].size() * 4, hipMemcpyHostToDevice, c->st));
        S->grid.cuts[a] = S->d_cuts + off;
        S->grid.ncut[a] = int(G.cuts[a].size());
        S->grid.dim[a] = G.dim[a];
        off += G.cuts[a].size();
    }
    HIPCK(hipMemcpyAsync(S->d_cell, G.cell.data(), G.cell.size() * 2, hipMemcpyHostToDevice, c->st));
    S->grid.cell = S->d_cell;
    S->clim_cap = 4096;
    HIPCK(hipMalloc(&S->d_clim, size_t(S->clim_cap) * sizeof(dsx::ClimSeed)));
    HIPCK(hipHostMalloc(&S->h_useeds, size_t(S->clim_cap) * 12, hipHostMallocDefault));
    // flood engine
    dsx::FloodBufs &B = S->fb;
    B.tile_cap = 1u << 21;                    // 2M tiles x sizeof(Tile) (672 B): 1.4 GB
    B.hmask = (1u << 23) - 1;                 // 8M hash slots
    HIPCK(hipMalloc(&B.tiles, size_t(B.tile_cap) * sizeof(dsx::Tile)));
    // every initialisation is ordered on the context's stream and completed before the state is used:
    // hipMemset would run on the null stream, which the non-blocking context streams do not wait for
    HIPCK(hipMemsetAsync(B.tiles, 0, size_t(B.tile_cap) * sizeof(dsx::Tile), c->st));
    HIPCK(hipMalloc(&B.hkeys, (size_t(B.hmask) + 1) * 8));
    HIPCK(hipMemsetAsync(B.hkeys, 0xff, (size_t(B.hmask) + 1) * 8, c->st));
    B.wq_cap = 4 * B.tile_cap;               // queue slots (a tile is queued again when touched again)
    HIPCK(hipMalloc(&B.wq, size_t(B.wq_cap) * 4));
    HIPCK(hipMemsetAsync(B.wq, 0xff, size_t(B.wq_cap) * 4, c->st));
    HIPCK(hipMalloc(&S->d_ctl, CTL_BLOCKS * dsx::CTL_WORDS * 4));
    HIPCK(hipMemsetAsync(S->d_ctl, 0, CTL_BLOCKS * dsx::CTL_WORDS * 4, c->st));
    S->job_cap = 1u << 20;
    HIPCK(hipMalloc(&S->d_up, size_t(S->job_cap) * (sizeof(dsx::FloodJob) + 8)));
    HIPCK(hipMalloc(&S->d_res, size_t(S->job_cap) * sizeof(dsx::FloodRes)));
    S->mask_cap = 1u << 20;
    HIPCK(hipMalloc(&S->d_masks, size_t(S->mask_cap) * sizeof(dsx::TileMask)));
    HIPCK(hipHostMalloc(&S->h_ctl, dsx::CTL_WORDS * 4, hipHostMallocDefault));
    HIPCK(hipStreamSynchronize(c->st));
    return DSX_OK;
}

// Pinned staging for a batch of n floods (grown between calls, when the stream is idle). The first
// sizes hold a pipeline worker's batches (65,536 floods and masks, ~14 MB pinned) and buffers at least
// double when they grow: hipFree / hipHostFree wait for the whole device, so every reallocation stalls
// the calling worker behind all other streams' work.
static int flood_io_reserve(ClimState *S, uint64_t n, uint64_t nh) {
    const size_t up = size_t(n) * (sizeof(dsx::FloodJob) + 8) + size_t(nh) * sizeof(dsx::FloodHint);
    if (S->h_up_cap < up) {
        if (S->h_up) HIPCK(hipHostFree(S->h_up));
        S->h_up = nullptr;
        S->h_up_cap = std::max({up, 2 * S->h_up_cap, size_t(1 << 16) * (sizeof(dsx::FloodJob) + 8)});
        HIPCK(hipHostMalloc(&S->h_up, S->h_up_cap, hipHostMallocDefault));
    }
    if (S->h_res_cap < n) {
        if (S->h_res) HIPCK(hipHostFree(S->h_res));
        S->h_res = nullptr;
        S->h_res_cap = uint32_t(std::max<uint64_t>({n, 2 * uint64_t(S->h_res_cap), 1u << 16}));
        HIPCK(hipHostMalloc(&S->h_res, size_t(S->h_res_cap) * sizeof(dsx::FloodRes), hipHostMallocDefault));
    }
    if (!S->h_masks || S->h_mask_cap < S->mask_want) {   // mask_want: the last overflow of the pinned masks
        if (S->h_masks) HIPCK(hipHostFree(S->h_masks));
        S->h_masks = nullptr;
        S->h_mask_cap = std::max(1u << 16, S->mask_want);
        HIPCK(hipHostMalloc(&S->h_masks, size_t(S->h_mask_cap) * sizeof(dsx::TileMask), hipHostMallocDefault));
    }
    return DSX_OK;
}

// Device I/O buffer and its pinned host twin (one upload and one download per call).
static int io_reserve(dsx_ctx *c, ClimState *S, size_t bytes) {
    if (S->io_cap >= bytes) return DSX_OK;
    if (S->d_io) HIPCK(hipFree(S->d_io));
    if (S->h_io) HIPCK(hipHostFree(S->h_io));
    S->d_io = nullptr;
    S->h_io = nullptr;
    S->io_cap = std::max({bytes, 2 * S->io_cap, size_t(8) << 20});   // at least doubles (see flood_io_reserve)
    HIPCK(hipMalloc(&S->d_io, S->io_cap));
    HIPCK(hipHostMalloc(&S->h_io, S->io_cap, hipHostMallocDefault));
    return DSX_OK;
}

// Climate slots of `n` seeds (repeats allowed): slot[k] = index into S->d_clim. A seed's climate stays
// on the device across calls; the seeds not cached are initialised in one k_climinit launch (when they
// do not fit the free slots, the cache restarts empty).
static int clim_slots(dsx_ctx *c, ClimState *S, const uint64_t *seeds, uint32_t n, uint32_t *slot) {
    std::vector<uint64_t> &miss = S->miss;
    auto assign = [&]() {
        miss.clear();
        for (uint32_t k = 0; k < n; k++) {
            auto it = S->cslot.find(seeds[k]);
            if (it != S->cslot.end()) {
                slot[k] = it->second;
                continue;
            }
            if (S->cslot.size() >= S->clim_cap) return false;
            slot[k] = uint32_t(S->cslot.size());
            S->cslot.emplace(seeds[k], slot[k]);
            miss.push_back(seeds[k]);
        }
        return true;
    };
    if (!assign()) {
        S->cslot.clear();   // restart the cache with this call's seeds
        if (!assign()) {
            S->cslot.clear();
            dsx::g_last_error = "too many distinct seeds in one call";
            return DSX_E_CAPACITY;
        }
    }
    if (!miss.empty()) {   // the kernel reads the pinned list in place (every API call ends synchronised,
        const uint32_t m = uint32_t(miss.size());   // so no earlier k_climinit still reads it)
        uint32_t *h_sl = reinterpret_cast<uint32_t *>(S->h_useeds + S->clim_cap);
        for (uint32_t k = 0; k < m; k++) {
            S->h_useeds[k] = miss[k];
            h_sl[k] = S->cslot[miss[k]];
        }
        dsx::launch_climinit(c->st, S->h_useeds, h_sl, m, S->d_clim);
        HIPCK(hipGetLastError());
    }
    return DSX_OK;
}

static uint64_t env_u32(const char *name, uint64_t def) {
    const char *v = std::getenv(name);
    return v && *v ? std::strtoull(v, nullptr, 10) : def;
}
static bool env_flag(const char *name, bool def) {
    const char *v = std::getenv(name);
    return v && *v ? std::atoi(v) != 0 : def;
}

// Process-wide flood-engine statistics (dsx_flood_stats_get) and the profiling switch (dsx_profile).
static std::atomic<int> g_prof_on{0};
static std::mutex g_stats_mu;
static dsx_flood_stats g_stats{};
static void stats_add(const dsx_flood_stats &a) {
    std::lock_guard<std::mutex> l(g_stats_mu);
    g_stats.calls += a.calls;
    g_stats.rounds += a.rounds;
    for (int k = 0; k < 4; k++) {
        g_stats.floods[k] += a.floods[k];
        g_stats.tiles[k] += a.tiles[k];
        g_stats.evaluated[k] += a.evaluated[k];
        g_stats.sampled_cold[k] += a.sampled_cold[k];
        g_stats.sampled_hot[k] += a.sampled_hot[k];
    }
    g_stats.small_floods += a.small_floods;
    g_stats.small_sampled += a.small_sampled;
    g_stats.eval_ms += a.eval_ms;
    g_stats.eval_launches += a.eval_launches;
    g_stats.eval_rounds += a.eval_rounds;
    g_stats.eval_needed += a.eval_needed;
    g_stats.wg_wait += a.wg_wait;
    g_stats.wg_eval += a.wg_eval;
    g_stats.wg_load += a.wg_load;
    g_stats.wg_life += a.wg_life;
    for (int k = 0; k < 4; k++)
        for (int b = 0; b < 6; b++) {
            g_stats.size_floods[k][b] += a.size_floods[k][b];
            g_stats.size_tiles[k][b] += a.size_tiles[k][b];
            g_stats.size_sampled[k][b] += a.size_sampled[k][b];
        }
}

// All per-context extension state lives here; freed by dsx_destroy via ctx_free_ext.
#include <mutex>
static std::mutex g_ext_mu;
static std::unordered_map<dsx_ctx *, ClimState *> g_ext;
static ClimState *clim_state(dsx_ctx *c) {
    std::lock_guard<std::mutex> l(g_ext_mu);
    ClimState *&s = g_ext[c];
    if (!s) s = new ClimState();
    return s;
}
void dsx_ctx_free_ext(dsx_ctx *c) {
    ClimState *s = nullptr;
    {
        std::lock_guard<std::mutex> l(g_ext_mu);
        auto it = g_ext.find(c);
        if (it == g_ext.end()) return;
        s = it->second;
        g_ext.erase(it);
    }
    for (void *p : {(void *)s->d_cuts, (void *)s->d_cell, (void *)s->d_clim, (void *)s->fb.tiles,
                    (void *)s->fb.hkeys, (void *)s->fb.wq, (void *)s->d_ctl, (void *)s->d_up,
                    (void *)s->d_res, s->d_io, (void *)s->d_masks})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)s->h_up, (void *)s->h_res, (void *)s->h_ctl, (void *)s->h_masks, s->h_io, (void *)s->h_useeds})
        if (p) (void)hipHostFree(p);
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    delete s;
}

static int noise_impl(dsx_ctx *c, uint64_t seed, int channel, const double *xyz, uint64_t n, double *out, int fp32) {
    if (!c || channel < 0 || channel > 5 || (n && (!xyz || !out))) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    ClimState *S;
    int rc = clim_alloc(c, S);
    if (rc) return rc;
    uint32_t sl = 0;
    if ((rc = clim_slots(c, S, &seed, 1, &sl))) return rc;
    if ((rc = io_reserve(c, S, n * 32))) return rc;
    double *d_pts = static_cast<double *>(S->d_io), *d_out = d_pts + 3 * n;
    HIPCK(hipMemcpyAsync(d_pts, xyz, n * 24, hipMemcpyHostToDevice, c->st));
    dsx::launch_noise(c->st, S->d_clim + sl, channel, d_pts, uint32_t(n), d_out, fp32);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, d_out, n * 8, hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    return DSX_OK;
}

extern "C" {

int dsx_noise(dsx_ctx *c, uint64_t seed, int channel, const double *xyz, uint64_t n, double *out) {
    return noise_impl(c, seed, channel, xyz, n, out, 0);
}

int dsx_noise_f32(dsx_ctx *c, uint64_t seed, int channel, const double *xyz, uint64_t n, double *out) {
    return noise_impl(c, seed, channel, xyz, n, out, 1);
}

int dsx_set_precision(dsx_ctx *c, int mode) {
    if (!c || mode < 0 || mode > 1) return DSX_E_ARG;
    clim_state(c)->prepass = mode;
    return DSX_OK;
}

int dsx_climate(dsx_ctx *c, uint64_t seed, const int32_t *xs, const int32_t *zs, uint64_t n, int64_t *q) {
    if (!c || (n && (!xs || !zs || !q))) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    ClimState *S;
    int rc = clim_alloc(c, S);
    if (rc) return rc;
    uint32_t sl = 0;
    if ((rc = clim_slots(c, S, &seed, 1, &sl))) return rc;
    if ((rc = io_reserve(c, S, n * 56))) return rc;
    int32_t *d_x = static_cast<int32_t *>(S->d_io), *d_z = d_x + n;
    int64_t *d_q = reinterpret_cast<int64_t *>(d_z + n + (n & 1));
    HIPCK(hipMemcpyAsync(d_x, xs, n * 4, hipMemcpyHostToDevice, c->st));
    HIPCK(hipMemcpyAsync(d_z, zs, n * 4, hipMemcpyHostToDevice, c->st));
    dsx::launch_climate(c->st, S->d_clim + sl, d_x, d_z, uint32_t(n), d_q);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(q, d_q, n * 48, hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    return DSX_OK;
}

int dsx_predicates(dsx_ctx *c, uint64_t seed, int kind, float cut, const int32_t *xs, const int32_t *zs, uint64_t n,
                   int8_t *out) {
    if (!c || kind < 0 || kind > 3 || (n && (!xs || !zs || !out))) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    ClimState *S;
    int rc = clim_alloc(c, S);
    if (rc) return rc;
    uint32_t sl = 0;
    if ((rc = clim_slots(c, S, &seed, 1, &sl))) return rc;
    if ((rc = io_reserve(c, S, n * 9))) return rc;
    int32_t *d_x = static_cast<int32_t *>(S->d_io), *d_z = d_x + n;
    int8_t *d_o = reinterpret_cast<int8_t *>(d_z + n);
    HIPCK(hipMemcpyAsync(d_x, xs, n * 4, hipMemcpyHostToDevice, c->st));
    HIPCK(hipMemcpyAsync(d_z, zs, n * 4, hipMemcpyHostToDevice, c->st));
    dsx::launch_pred(c->st, S->d_clim + sl, S->grid, kind, cut, d_x, d_z, uint32_t(n), d_o, S->prepass);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, d_o, n, hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    return DSX_OK;
}

int dsx_flood(dsx_ctx *c, const dsx_flood_job *jobs, uint64_t n, dsx_area *out) {
    return dsx_flood_hinted(c, jobs, n, nullptr, 0, out);
}

int dsx_flood_hinted(dsx_ctx *c, const dsx_flood_job *jobs, uint64_t n, const dsx_flood_hint *hints, uint64_t nh,
                     dsx_area *out) {
    if (!c || (n && (!jobs || !out)) || (nh && !hints) || nh > (1u << 24)) return DSX_E_ARG;
    for (uint64_t k = 0; k < nh; k++)
        if (hints[k].job >= n) return DSX_E_ARG;
    for (uint64_t k = 0; k < n; k++) {
        const dsx_flood_job &j = jobs[k];
        if (j.kind < 0 || j.kind > 3 || j.step < 1 || j.half < 0 || 2 * int64_t(j.half) + 1 > DSX_FLOOD_MAX_WIDTH ||
            int64_t(j.half) * j.step > 1000000 || j.stop_cells < 0 || j.stop_cells > (1 << 23))
            return DSX_E_ARG;
    }
    if (n == 0) {
        ClimState *S0 = clim_state(c);
        S0->masks_n = S0->masks_h = 0;
        return DSX_OK;
    }
    HIPCK(hipSetDevice(c->device));
    ClimState *S;
    int rc = clim_alloc(c, S);
    if (rc) return rc;
    S->masks_n = S->masks_h = 0;   // until this call's masks are published
    if (n > S->job_cap) {
        dsx::g_last_error = "too many floods in one call";
        return DSX_E_CAPACITY;
    }
    if ((rc = flood_io_reserve(S, n, nh))) return rc;
    // jobs (seed -> climate slot) and job lists straight into pinned staging, in the device layout
    dsx::FloodJob *hj = reinterpret_cast<dsx::FloodJob *>(S->h_up);
    uint32_t *h_small = reinterpret_cast<uint32_t *>(hj + n), *h_big = h_small + n;
    std::vector<uint64_t> useeds;
    std::unordered_map<uint64_t, int32_t> slot;
    for (uint64_t k = 0; k < n; k++) {
        auto [it, fresh] = slot.try_emplace(jobs[k].seed, int32_t(useeds.size()));
        if (fresh) useeds.push_back(jobs[k].seed);
        hj[k] = dsx::FloodJob{it->second, jobs[k].kind, jobs[k].cut, jobs[k].qx, jobs[k].qz, jobs[k].step, jobs[k].half,
                              (jobs[k].conn8 != 0 ? 1 : 0) | (jobs[k].stop_cells << 8)};
    }
    std::vector<uint32_t> uslot(useeds.size());
    if ((rc = clim_slots(c, S, useeds.data(), uint32_t(useeds.size()), uslot.data()))) return rc;
    for (uint64_t k = 0; k < n; k++) hj[k].seed = int32_t(uslot[hj[k].seed]);
    // The one-wave LDS engine for every flood whose window fits it whole and, with DSX_LDS_ROUTE=1,
    // first for hot-zone and proxy floods of any window (on a local sub-window; exact work, but one
    // BFS layer per step, which measured slower in the pipeline than the tiled engine); the rest, and
    // the floods that outgrow the sub-window (handed over on the device), on the tiled engine.
    static const bool lds_route = env_flag("DSX_LDS_ROUTE", false);
    uint32_t n_cls[dsx::LDS_NCLS] = {}, n_small = 0, n_big = 0;
    static const bool hot_small = env_flag("DSX_LDS_HOT_SMALL", true);
    // proxy floods of any window on a 64 x 64 sub-window first (most touch 1-4 tiles)
    static const bool proxy_small = env_flag("DSX_LDS_PROXY_SMALL", false);
    auto cls_of = [&](const dsx::FloodJob &j) {
        const bool fits = 2 * j.half + 1 <= dsx::LDS_MAX_N;
        // threshold floods (a stop count: the hot-zone pre-floods) almost always end in a few dozen
        // cells: a 64 x 64 sub-window (7 KB of LDS instead of 28) holds them; the rest are handed over
        if (j.kind == dsx::PRED_HOT && (uint32_t(j.conn8) >> 8) != 0u && hot_small && (fits || lds_route))
            return int(dsx::LDS_HOT_SMALL);
        if (j.kind == dsx::PRED_HOT && (fits || lds_route)) return int(dsx::LDS_HOT);
        if (j.kind == dsx::PRED_PROXY && proxy_small) return int(dsx::LDS_PROXY_SMALL);
        if (j.kind == dsx::PRED_PROXY && (fits || lds_route)) return int(dsx::LDS_PROXY);
        return fits ? int(dsx::LDS_FULL) : -1;
    };
    for (int cls = 0; cls < dsx::LDS_NCLS; cls++)
        for (uint64_t k = 0; k < n; k++)
            if (cls_of(hj[k]) == cls) {
                h_small[n_small++] = uint32_t(k);
                n_cls[cls]++;
            }
    // every tiled job a hot-zone or proxy flood, including those the small-window engine may hand over
    // (any of its floods whose ring overflows): so no small-window biome-predicate floods either
    bool light = n_cls[dsx::LDS_FULL] == 0;
    for (uint64_t k = 0; k < n; k++)
        if (cls_of(hj[k]) < 0) {
            h_big[n_big++] = uint32_t(k);
            light = light && (hj[k].kind == dsx::PRED_HOT || hj[k].kind == dsx::PRED_PROXY);
        }
    // hints of the tiled engine's floods, after the job lists (the small-window engine takes no hints)
    dsx::FloodHint *h_hint = reinterpret_cast<dsx::FloodHint *>(h_small + 2 * n);
    uint32_t n_hint = 0;
    for (uint64_t k = 0; k < nh; k++)
        if (cls_of(hj[hints[k].job]) < 0 && (hints[k].blocks & 0xffffu))
            h_hint[n_hint++] = dsx::FloodHint{hints[k].job, hints[k].tx, hints[k].tz, hints[k].blocks & 0xffffu};
    uint8_t *d_up = S->d_up;
    const dsx::FloodJob *d_jobs = reinterpret_cast<const dsx::FloodJob *>(d_up);
    uint32_t *d_small = reinterpret_cast<uint32_t *>(d_up + n * sizeof(dsx::FloodJob));
    dsx::FloodBufs &B = S->fb;
    // a batch that reaches the engine's time limit is run once more on a fully reset engine (tiles, hash
    // table, work queue and both control blocks); its results are only taken from a complete run
    const uint32_t *h = S->h_ctl;
    const dsx::FloodRes *res = S->h_res;
    const bool prof = g_prof_on.load(std::memory_order_relaxed) != 0;
    const bool tiled = n_big > 0 || n_small > 0;
    uint32_t blocks_used = 0;
    for (int attempt = 0;; attempt++) {
        // this batch's control block was zeroed by the publish kernel of the batch two before it; this
        // batch's publish zeroes the block of the batch two after it
        const uint32_t seq = S->ctl_seq++;
        B.ctl = S->d_ctl + (seq % CTL_BLOCKS) * dsx::CTL_WORDS;
        B.ctl_next = S->d_ctl + ((seq + 2) % CTL_BLOCKS) * dsx::CTL_WORDS;
        B.ntiles = B.ctl + dsx::CTL_NTILES;
        B.seq = seq;
        // with the fused accumulate-and-publish kernel, masks reach pinned memory as they are exported
        static const bool acc_publish = env_flag("DSX_FLOOD_ACC_PUBLISH", true);
        const bool fused = tiled && acc_publish;
        B.h_masks = fused ? S->h_masks : nullptr;
        B.h_mcap = fused ? S->h_mask_cap : 0u;
        // jobs and lists read from the pinned staging by the first kernel (no separate copy): the stage
        // kernel, or (a batch without small-window floods) the tiled engine's own first workgroups
        static const bool self_stage = env_flag("DSX_FLOOD_SELF_STAGE", true);
        const bool staged_by_run = self_stage && n_small == 0 && n_big > 0;
        if (!staged_by_run)
            dsx::launch_flood_stage(c->st, S->h_up, uint32_t(n), n_small, n_big, n_hint, d_up, S->d_res, B);
        uint32_t off = 0;
        for (int cls = 0; cls < dsx::LDS_NCLS; cls++) {
            dsx::launch_flood_lds(c->st, cls, d_jobs, d_small + off, n_cls[cls], S->d_clim, S->grid, S->d_res,
                                  S->d_masks, S->mask_cap, B, S->prepass);
            off += n_cls[cls];
        }
        // the tiled engine: the jobs listed now plus the small-window floods handed over on the device
        if (tiled) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (prof) {
                while (S->events.size() < 2) {
                    hipEvent_t e;
                    HIPCK(hipEventCreate(&e));
                    S->events.push_back(e);
                }
                e0 = S->events[0];
                e1 = S->events[1];
            }
            // workgroups (in units of 256 threads): enough for the tiles in flight. Hot-zone and proxy
            // floods are short chains of cheap tiles; exact floods (52 octave samples per hot cell) keep
            // their workgroups evaluating most of the time, so a batch of them gets a wider launch
            static const uint32_t per_job = uint32_t(env_u32("DSX_FLOOD_BLOCKS_PER_JOB", 4)),
                                  max_blocks = uint32_t(env_u32("DSX_FLOOD_MAX_BLOCKS", 128)),
                                  min_full = uint32_t(env_u32("DSX_FLOOD_MIN_BLOCKS_FULL", 16)),
                                  per_full = uint32_t(env_u32("DSX_FLOOD_BLOCKS_PER_EXACT", 4)),
                                  max_full = uint32_t(env_u32("DSX_FLOOD_MAX_BLOCKS_FULL", 128));
            // hinted tiles are evaluated in bulk at the start: one more unit per `per_hint` of them
            static const uint32_t per_hint = std::max(1u, uint32_t(env_u32("DSX_FLOOD_TILES_PER_UNIT", 8)));
            const size_t want = (light ? size_t(per_job) : size_t(per_full)) * n_big + n_hint / per_hint;
            const uint32_t blocks =
                light ? uint32_t(std::clamp<size_t>(want, 16, max_blocks))
                      : uint32_t(std::clamp<size_t>(want, std::min(min_full, max_full), max_full));
            blocks_used = blocks;
            const dsx::StageArgs sa{S->h_up, d_up, S->d_res, uint32_t(n), n_big, n_hint, 0u};
            dsx::launch_flood_run(c->st, d_jobs, S->d_clim, S->grid, B, blocks, S->prepass, light ? 1 : 0, e0, e1,
                                  prof ? 1 : 0, staged_by_run ? &sa : nullptr);
        }
        // accumulate, then publish: one kernel whose last workgroup publishes (DSX_FLOOD_ACC_PUBLISH=0: two)
        if (fused) {
            dsx::launch_flood_acc_publish(c->st, d_jobs, B, S->d_res, S->d_masks, S->mask_cap, n_big ? 1024 : 64,
                                          uint32_t(n), S->h_res, S->h_ctl, S->h_masks, S->h_mask_cap);
        } else {
            if (tiled) dsx::launch_flood_acc(c->st, d_jobs, B, S->d_res, S->d_masks, S->mask_cap, n_big ? 1024 : 64);
            dsx::launch_flood_publish(c->st, S->d_res, uint32_t(n), B, S->d_masks, S->h_res, S->h_ctl, S->h_masks,
                                      S->h_mask_cap);
        }
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->st));
        if (!h[dsx::FLOOD_OVF]) break;
        // acc ran on an incomplete flood state: results are void, and the engine state is reset
        int status = DSX_E_CAPACITY;
        dsx::g_last_error = "flood tile pool or work queue exhausted: split the batch";
        const bool timed_out = (h[dsx::FLOOD_OVF] & dsx::OVF_TIME) != 0;
        if (timed_out) {
            status = DSX_E_HIP;
            char msg[384];
            std::snprintf(msg, sizeof msg,
                          "flood engine time limit reached (slot %u: head %u tail %u pending %u item %08x; "
                          "%u floods, %u tiled, %u workgroups; batch %u: stage %u, engine %u saw tail %u stage %u)",
                          h[dsx::TIME_SLOT], h[dsx::TIME_HEAD], h[dsx::TIME_TAIL], h[dsx::TIME_PENDING],
                          h[dsx::TIME_ITEM], unsigned(n), n_big, blocks_used, seq + 1, h[dsx::STAGE_SEQ],
                          h[dsx::RUN_SEQ], h[dsx::RUN_TAIL0], h[dsx::RUN_STAGE_SEQ]);
            dsx::g_last_error = msg;
        }
        {
            std::lock_guard<std::mutex> l(g_stats_mu);
            if (timed_out) g_stats.time_retries++;
            else g_stats.capacity_resets++;
        }
        const uint32_t ntiles = timed_out ? B.tile_cap : std::min(h[dsx::CTL_NTILES], B.tile_cap);
        const uint32_t qtail = timed_out ? B.wq_cap : std::min(h[dsx::Q_TAIL], B.wq_cap);
        HIPCK(hipMemsetAsync(B.tiles, 0, size_t(ntiles) * sizeof(dsx::Tile), c->st));
        HIPCK(hipMemsetAsync(B.hkeys, 0xff, (size_t(B.hmask) + 1) * 8, c->st));
        if (qtail) HIPCK(hipMemsetAsync(B.wq, 0xff, size_t(qtail) * 4, c->st));
        if (timed_out) HIPCK(hipMemsetAsync(S->d_ctl, 0, CTL_BLOCKS * dsx::CTL_WORDS * 4, c->st));
        HIPCK(hipStreamSynchronize(c->st));
        if (!timed_out || attempt > 0) return status;
        std::fprintf(stderr, "dsx_flood: %s; running the batch again on a reset engine\n", dsx::g_last_error.c_str());
    }
    const uint32_t nm = h[dsx::FLOOD_MASKS];
    if (nm > S->mask_cap) {
        dsx::g_last_error = "flood mask export overflow";
        return DSX_E_CAPACITY;
    }
    // the masks stay in the pinned buffer the publish kernel wrote (dsx_flood_masks copies them out); only
    // those beyond its capacity are copied here
    const uint32_t nmh = std::min(nm, S->h_mask_cap);
    S->masks.resize(nm - nmh);
    if (nm > nmh) {   // the next call's staging holds this many pinned (twice the current capacity at least)
        S->mask_want = std::min(S->mask_cap, std::max(nm, 2 * S->h_mask_cap));
        HIPCK(hipMemcpyAsync(S->masks.data(), S->d_masks + nmh, size_t(nm - nmh) * sizeof(dsx::TileMask),
                             hipMemcpyDeviceToHost, c->st));
        HIPCK(hipStreamSynchronize(c->st));
    }
    S->masks_n = nm;
    S->masks_h = nmh;
    dsx_flood_stats st{};
    st.calls = 1;
    st.rounds = h[dsx::Q_ITEMS];
    st.eval_rounds = h[dsx::EVAL_ROUNDS];
    st.eval_needed = h[dsx::EVAL_NEED];
    st.wg_wait = h[dsx::PROF_WAIT];
    st.wg_eval = h[dsx::PROF_EVAL];
    st.wg_load = h[dsx::PROF_LOAD];
    st.wg_life = h[dsx::PROF_LIFE];
    for (uint64_t k = 0; k < n; k++) {
        const int kind = hj[k].kind;
        if (res[k].tiles) {   // ran on the tiled engine
            st.floods[kind]++;
            st.tiles[kind] += res[k].tiles;
            st.evaluated[kind] += res[k].evald;
            st.sampled_hot[kind] += res[k].hot;
            st.sampled_cold[kind] += res[k].sampled - res[k].hot;
            int b = 0;
            while (b < 5 && res[k].tiles > (1ull << (2 * b))) b++;
            st.size_floods[kind][b]++;
            st.size_tiles[kind][b] += res[k].tiles;
            st.size_sampled[kind][b] += res[k].sampled;
        } else {
            st.small_floods++;
            st.small_sampled += res[k].sampled;
        }
    }
    if (prof && tiled) {
        float ms = 0;
        HIPCK(hipEventElapsedTime(&ms, S->events[0], S->events[1]));
        st.eval_ms += ms;
        st.eval_launches++;
    }
    stats_add(st);
    for (uint64_t k = 0; k < n; k++)
        out[k] = dsx_area{res[k].cells, res[k].sampled, res[k].minx, res[k].maxx, res[k].minz, res[k].maxz, res[k].edge,
                          res[k].border};
    return DSX_OK;
}

int dsx_profile(int on) {
    g_prof_on.store(on ? 1 : 0);
    return DSX_OK;
}

int dsx_flood_stats_get(dsx_flood_stats *out, int reset) {
    if (!out) return DSX_E_ARG;
    std::lock_guard<std::mutex> l(g_stats_mu);
    *out = g_stats;
    if (reset) g_stats = dsx_flood_stats{};
    return DSX_OK;
}

int dsx_flood_masks(dsx_ctx *c, dsx_tile_mask *out, uint64_t cap, uint64_t *n_out) {
    if (!c || !n_out || (cap && !out)) return DSX_E_ARG;
    ClimState *S = clim_state(c);
    *n_out = S->masks_n;
    static_assert(sizeof(dsx_tile_mask) == sizeof(dsx::TileMask), "mask layout");
    const uint64_t w = std::min<uint64_t>(cap, S->masks_n), wh = std::min<uint64_t>(w, S->masks_h);
    if (wh) std::memcpy(out, S->h_masks, wh * sizeof(dsx_tile_mask));
    if (w > wh) std::memcpy(out + wh, S->masks.data(), (w - wh) * sizeof(dsx_tile_mask));
    return S->masks_n > cap ? DSX_E_CAPACITY : DSX_OK;
}

}  // extern "C"

namespace dsxh {
// The last dsx_flood call's masks where they lie (no copy): a (pinned memory), then b (the overflow
// beyond the pinned buffer). Valid until the context's next flood call.
int flood_masks_view(dsx_ctx *c, const dsx_tile_mask **a, uint64_t *na, const dsx_tile_mask **b, uint64_t *nb) {
    if (!c || !a || !na || !b || !nb) return DSX_E_ARG;
    ClimState *S = clim_state(c);
    static_assert(sizeof(dsx_tile_mask) == sizeof(dsx::TileMask), "mask layout");
    *na = S->masks_h;
    *a = reinterpret_cast<const dsx_tile_mask *>(S->h_masks);
    *nb = S->masks_n - S->masks_h;
    *b = reinterpret_cast<const dsx_tile_mask *>(S->masks.data());
    return DSX_OK;
}
}  // namespace dsxh

extern "C" {

int dsx_points(dsx_ctx *c, const uint64_t *seeds, uint64_t nseeds, const uint32_t *sidx, const int32_t *xs,
               const int32_t *zs, uint64_t n, int what, float cut, int8_t *out8, double *outd) {
    const bool hottest = what == DSX_POINT_HOTTEST;
    const bool dbl = what == DSX_POINT_T || what == DSX_POINT_PROXYPROB || what == DSX_POINT_PROXY_C ||
                     what == DSX_POINT_PROXY_E || hottest;
    if (!c || (n && (!sidx || !xs || !zs)) || (!dbl && (what < 0 || what > 3) && what != DSX_POINT_BIOME) || (dbl && n && !outd) ||
        ((!dbl || hottest) && n && !out8) || (nseeds && !seeds))
        return DSX_E_ARG;
    for (uint64_t k = 0; k < n; k++)
        if (sidx[k] >= nseeds) return DSX_E_ARG;
    if (n == 0) return DSX_OK;
    HIPCK(hipSetDevice(c->device));
    ClimState *S;
    int rc = clim_alloc(c, S);
    if (rc) return rc;
    if (nseeds > S->clim_cap) {
        dsx::g_last_error = "too many distinct seeds in one call";
        return DSX_E_CAPACITY;
    }
    std::vector<uint32_t> sslot(nseeds), sidx2(n);
    if ((rc = clim_slots(c, S, seeds, uint32_t(nseeds), sslot.data()))) return rc;
    for (uint64_t k = 0; k < n; k++) sidx2[k] = sslot[sidx[k]];
    if ((rc = io_reserve(c, S, n * 24 + 16))) return rc;
    // inputs packed in pinned host memory, read by the kernel in place (each once, coalesced), and
    // results written by the kernel straight after them (no copy dispatches):
    // [sidx | xs | zs | pad] then the doubles (or bytes), then for HOTTEST the index bytes
    const size_t in_bytes = size_t(n) * 12 + (n & 1) * 4;
    uint32_t *h_s = static_cast<uint32_t *>(S->h_io);
    std::memcpy(h_s, sidx2.data(), n * 4);
    std::memcpy(h_s + n, xs, n * 4);
    std::memcpy(h_s + 2 * n, zs, n * 4);
    const int32_t *h_x = reinterpret_cast<const int32_t *>(h_s + n), *h_z = h_x + n;
    uint8_t *h_out = static_cast<uint8_t *>(S->h_io) + in_bytes;
    double *h_d = reinterpret_cast<double *>(h_out);
    int8_t *h_8 = hottest ? reinterpret_cast<int8_t *>(h_d + n) : reinterpret_cast<int8_t *>(h_d);
    dsx::launch_points(c->st, S->d_clim, S->grid, h_s, h_x, h_z, uint32_t(n), what, cut, h_8, h_d, S->prepass);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->st));
    if (dbl) std::memcpy(outd, h_out, n * 8);
    if (hottest) std::memcpy(out8, h_out + n * 8, n);
    else if (!dbl) std::memcpy(out8, h_out, n);
    return DSX_OK;
}

int dsx_dfunc(dsx_ctx *c, uint32_t nseeds, uint32_t pts, uint64_t *hist) {
    if (!c || !hist || pts > 10000000u) return DSX_E_ARG;
    std::memset(hist, 0, DSX_DFUNC_CELLS * sizeof(uint64_t));
    // seeds mix(7k + 99991) that pass the temperature phase filter s <= 0.36 (tools/dfunc.cpp:37-45)
    std::vector<uint64_t> seeds;
    for (uint32_t k = 0; k < nseeds; k++) {
        const uint64_t seed = dsx::mix64(uint64_t(k) * 7 + 99991);
        if (!(dsx::temperature_phase(seed, nullptr) > 0.36)) seeds.push_back(seed);
    }
    if (seeds.empty() || pts == 0) return DSX_OK;
    HIPCK(hipSetDevice(c->device));
    ClimState *S;
    int rc = clim_alloc(c, S);
    if (rc) return rc;
    // io layout: histogram (u64 cells) | tries: sidx, x, z (u32 x T each) | hot flags (T bytes) |
    // selected hot points: sidx, x, z (u32 x T each)
    const uint32_t T = 1u << 22;
    const size_t o_try = DSX_DFUNC_CELLS * 8, o_hot = o_try + size_t(T) * 12, o_rec = o_hot + T;
    if ((rc = io_reserve(c, S, o_rec + size_t(T) * 12))) return rc;
    uint8_t *d = static_cast<uint8_t *>(S->d_io), *h = static_cast<uint8_t *>(S->h_io);
    unsigned long long *d_hist = reinterpret_cast<unsigned long long *>(d);
    HIPCK(hipMemsetAsync(d_hist, 0, DSX_DFUNC_CELLS * 8, c->st));
    uint32_t *h_try = reinterpret_cast<uint32_t *>(h + o_try), *h_rec = reinterpret_cast<uint32_t *>(h + o_rec);
    const uint32_t *d_try = reinterpret_cast<const uint32_t *>(d + o_try),
                   *d_rec = reinterpret_cast<const uint32_t *>(d + o_rec);
    const uint8_t *h_hot = h + o_hot;
    const uint64_t max_tries = uint64_t(pts) * 200;   // tools/dfunc.cpp:50
    struct Run {
        dsx::Xoro rng;
        uint64_t tries;
        uint32_t got, slot;
        uint32_t t0, m;   // this round's tries in the batch
    };
    const uint32_t G = 1024;   // seeds per group (climate slots)
    for (size_t g0 = 0; g0 < seeds.size(); g0 += G) {
        const uint32_t ng = uint32_t(std::min<size_t>(G, seeds.size() - g0));
        std::vector<uint32_t> slot(ng);
        if ((rc = clim_slots(c, S, seeds.data() + g0, ng, slot.data()))) return rc;
        std::vector<Run> run(ng);
        std::vector<uint32_t> active(ng);
        for (uint32_t i = 0; i < ng; i++) {
            run[i] = {dsx::Xoro::seeded(seeds[g0 + i] ^ 0x1234), 0, 0, slot[i], 0, 0};
            active[i] = i;
        }
        while (!active.empty()) {
            // this round: each active seed draws up to 8x the hot points it still needs (at least 256), in
            // draw order, while the batch has room; the batch is laid out sidx[n] | x[n] | z[n]
            uint32_t n = 0;
            std::vector<uint32_t> in_round;
            for (uint32_t i : active) {
                Run &r = run[i];
                const uint64_t want = std::max<uint64_t>(256, uint64_t(pts - r.got) * 8);
                const uint32_t m = uint32_t(std::min<uint64_t>(std::min(want, max_tries - r.tries), T - n));
                if (m == 0) break;
                r.t0 = n;
                r.m = m;
                n += m;
                in_round.push_back(i);
            }
            uint32_t *hs = h_try, *hx = h_try + n, *hz = h_try + 2 * size_t(n);
            for (uint32_t i : in_round) {
                Run &r = run[i];
                for (uint32_t t = r.t0; t < r.t0 + r.m; t++) {
                    hs[t] = r.slot;
                    hx[t] = uint32_t(int32_t(r.rng.bounded(15000000)) - 7500000);
                    hz[t] = uint32_t(int32_t(r.rng.bounded(15000000)) - 7500000);
                }
            }
            HIPCK(hipMemcpyAsync(const_cast<uint32_t *>(d_try), h_try, size_t(n) * 12, hipMemcpyHostToDevice, c->st));
            dsx::launch_dfunc_hot(c->st, S->d_clim, d_try, reinterpret_cast<const int32_t *>(d_try + n),
                                  reinterpret_cast<const int32_t *>(d_try + 2 * size_t(n)), n, d + o_hot);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(h + o_hot, d + o_hot, n, hipMemcpyDeviceToHost, c->st));
            HIPCK(hipStreamSynchronize(c->st));
            // the first hot tries of each seed in draw order, up to pts in all (tools/dfunc.cpp:50-56)
            std::vector<uint32_t> rs, rx, rz;
            for (uint32_t i : in_round) {
                Run &r = run[i];
                for (uint32_t t = r.t0; t < r.t0 + r.m && r.got < pts; t++) {
                    r.tries++;
                    if (!h_hot[t]) continue;
                    r.got++;
                    rs.push_back(r.slot);
                    rx.push_back(hx[t]);
                    rz.push_back(hz[t]);
                }
            }
            std::vector<uint32_t> still;
            for (uint32_t i : active)
                if (run[i].got < pts && run[i].tries < max_tries) still.push_back(i);
            active.swap(still);
            const uint32_t nr = uint32_t(rs.size());
            if (nr) {
                std::memcpy(h_rec, rs.data(), size_t(nr) * 4);
                std::memcpy(h_rec + nr, rx.data(), size_t(nr) * 4);
                std::memcpy(h_rec + 2 * size_t(nr), rz.data(), size_t(nr) * 4);
                HIPCK(hipMemcpyAsync(const_cast<uint32_t *>(d_rec), h_rec, size_t(nr) * 12, hipMemcpyHostToDevice,
                                     c->st));
                dsx::launch_dfunc_rec(c->st, S->d_clim, S->grid, d_rec, reinterpret_cast<const int32_t *>(d_rec + nr),
                                      reinterpret_cast<const int32_t *>(d_rec + 2 * size_t(nr)), nr, d_hist);
                HIPCK(hipGetLastError());
            }
        }
    }
    HIPCK(hipMemcpyAsync(h, d_hist, DSX_DFUNC_CELLS * 8, hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    std::memcpy(hist, h, DSX_DFUNC_CELLS * 8);
    return DSX_OK;
}

}  // extern "C"
