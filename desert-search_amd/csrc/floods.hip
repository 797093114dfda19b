// Tiled frontier flood engine for gfx950 (bounded bridged floods, dense_flood.hpp:52-86).
//
// The reference floods depth-first through a dense (2*half+1)^2 byte window. Every output of that
// flood (cells, sampled, bbox, edge, border, and the final window state) depends only on the set of
// cells reachable from the start through predicate != 0 cells and on their neighbourhood, not on the
// visiting order. This engine computes the same sets tile-parallel:
//   * the window is split into 32x32 tiles, allocated on first touch through a device hash table;
//   * k_flood_eval evaluates the FP64 predicate for all 1024 cells of each new tile (one 1024-thread
//     block per tile, one cell per thread, the octaves the predicate reads in LDS);
//   * k_flood_prop (one wave per tile, one lane per 32-bit row mask) grows the reached set inside the
//     tile to a fixpoint with bit-parallel dilation, then ORs the touched cells it adds across the
//     tile edges into the neighbours' `in` masks and queues those tiles;
//   * the host queues rounds of eval + prop until no tile is queued; k_flood_acc then reduces each tile:
//       reached = pushed cells, touched = reached | in | dilate_inside(reached),
//       cells = |reached & pred==2|, sampled = |touched inside the world|, border = touched outside
//       the world, edge = reached cell on the window boundary, bbox over counted reached cells
//       (starting at the window centre, as dense_flood.hpp:57).
#include "floods.h"
#include <hip/hip_ext.h>
#include <climits>

namespace dsx {

__device__ __forceinline__ uint32_t nwin(const FloodJob &J) { return uint32_t(2 * J.half + 1); }
__device__ __forceinline__ uint32_t ntile_axis(const FloodJob &J) { return (nwin(J) + FT - 1) / FT; }
// valid-column mask of a tile row: columns whose window index is < n
__device__ __forceinline__ uint32_t col_mask(const FloodJob &J, int tx) {
    const int n = int(nwin(J)), c0 = tx * FT;
    const int k = n - c0;
    return k >= FT ? 0xffffffffu : (k <= 0 ? 0u : ((1u << k) - 1u));
}

constexpr unsigned long long HEMPTY = ~0ull;
__device__ __forceinline__ unsigned long long hkey(uint32_t job, int tx, int tz) {
    return (uint64_t(job) << 44) | (uint64_t(uint32_t(tx) & 2047u) << 33) | (uint64_t(uint32_t(tz) & 2047u) << 22);
}
// Returns the tile index of (job, tx, tz), allocating it if absent; *fresh = this call allocated it.
// Keys and values are published together in one 64-bit CAS, so no reader ever sees a half-made entry.
__device__ uint32_t tile_get(const FloodBufs &B, uint32_t job, int tx, int tz, bool *fresh) {
    const unsigned long long key = hkey(job, tx, tz);
    uint32_t h = uint32_t(mix64(key) & B.hmask);
    uint32_t mine = 0xffffffffu;
    *fresh = false;
    for (uint32_t probe = 0; probe <= B.hmask; probe++, h = (h + 1) & B.hmask) {
        unsigned long long cur = __atomic_load_n(&B.hkeys[h], __ATOMIC_RELAXED);
        if (cur == HEMPTY) {
            if (mine == 0xffffffffu) {
                mine = atomicAdd(B.ntiles, 1u);
                if (mine >= B.tile_cap) return 0xffffffffu;        // pool exhausted (host reports it)
                Tile &t = B.tiles[mine];
                t.job = job;
                t.tx = tx;
                t.tz = tz;
                __threadfence();
            }
            const unsigned long long want = key | mine;
            cur = atomicCAS(&B.hkeys[h], HEMPTY, want);
            if (cur == HEMPTY) {
                B.tiles[mine].slot1 = h + 1;
                *fresh = true;
                return mine;
            }
        }
        if ((cur & ~0x3fffffull) == key) return uint32_t(cur & 0x3fffffull);
    }
    return 0xffffffffu;
}
// Octave slots read by the hot test (first 4: T of both sides) and the proxy (all 14: plus C octaves
// 0-2 and E octaves 0-1 of both sides).
__device__ __forceinline__ int proxy_slot(int k) {
    constexpr int slots[14] = {0, 1, 2, 3, 8, 9, 10, 17, 18, 19, 26, 27, 30, 31};
    return slots[k];
}
__device__ __forceinline__ void list_push(uint32_t *list, uint32_t *count, uint32_t cap, uint32_t *ovf, uint32_t v) {
    const uint32_t k = atomicAdd(count, 1u);
    if (k < cap) list[k] = v;
    else atomicOr(ovf, 1u);
}

__global__ void k_flood_resinit(const FloodJob *jobs, const uint32_t *jl, uint32_t n, FloodRes *res) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t k = jl ? jl[q] : q;
    FloodRes r{};
    r.minx = r.maxx = jobs[k].qx;
    r.minz = r.maxz = jobs[k].qz;
    res[k] = r;
}

// Appends v to list for every lane with on set: one atomic per wave.
__device__ __forceinline__ void wave_push(uint32_t *list, uint32_t *count, uint32_t cap, uint32_t *ovf, bool on, uint32_t v) {
    const uint64_t mask = __ballot(on);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, uint32_t(__popcll(mask)));
    base = __shfl(base, 0);
    if (on) {
        const uint32_t k = base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
        if (k < cap) list[k] = v;
        else atomicOr(ovf, 1u);
    }
}

// The start cell of every flood: touched ("in") and queued.
__global__ __launch_bounds__(256) void k_flood_start(const FloodJob *jobs, const uint32_t *jl, uint32_t n, FloodBufs B,
                                                     uint32_t *eval, uint32_t *prop) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    bool fresh = false, queue = false;
    uint32_t ti = 0xffffffffu;
    if (q < n) {
        const uint32_t k = jl[q];
        const FloodJob J = jobs[k];
        const int c = J.half % FT, t = J.half / FT;   // the centre cell (half, half)
        ti = tile_get(B, k, t, t, &fresh);
        if (ti == 0xffffffffu) {
            atomicOr(&B.lcount[FLOOD_OVF], 2u);
        } else {
            atomicOr(&B.tiles[ti].in[c], 1u << c);
            queue = atomicCAS(&B.tiles[ti].queued, 0u, 1u) == 0u;
        }
    }
    wave_push(eval, &B.lcount[0], B.list_cap, &B.lcount[FLOOD_OVF], fresh && ti != 0xffffffffu, ti);
    wave_push(prop, &B.lcount[1], B.list_cap, &B.lcount[FLOOD_OVF], queue, ti);
}

// Rounds: round r evaluates the new tiles of eval list r % 3 and propagates the tiles of prop list
// r % 3, listing the next round's tiles in lists (r + 1) % 3; its eval kernel first empties lists
// (r + 2) % 3 (read by the previous round, written by the next). The grids are fixed (grid-stride over
// the device-side counts), so the host queues several rounds per synchronisation.
//
// Predicate of the 1024 cells of each listed tile: pass = v != 0, cnt = v == 2 (cells outside the window
// or the world are neither).
__global__ __launch_bounds__(FT * FT) void k_flood_eval(const FloodJob *jobs, const ClimSeed *cs, SurfGrid g, FloodBufs B,
                                                        int round) {
    __shared__ ClimSeed C;
    const int tid = threadIdx.x;
    const int cur = round % 3, clr = (round + 2) % 3;
    const uint32_t n = min(__atomic_load_n(&B.lcount[2 * cur], __ATOMIC_RELAXED), B.list_cap);
    const uint32_t *list = B.lists[2 * cur];
    if (blockIdx.x == 0 && tid < 2) B.lcount[2 * clr + tid] = 0u;
    int loaded_seed = -1, loaded_cls = -1;
    for (uint32_t b = blockIdx.x; b < n; b += gridDim.x) {
        Tile &T = B.tiles[list[b]];
        const FloodJob J = jobs[T.job];
        // octaves the predicate reads: T for the hot test (class 0), T + C octaves 0-2 + E octaves 0-1
        // for the proxy (class 1), everything for the biome predicates (class 2)
        const int cls = J.kind == PRED_HOT ? 0 : (J.kind == PRED_PROXY ? 1 : 2);
        if (J.seed != loaded_seed || cls > loaded_cls) {
            __syncthreads();
            const ClimSeed &G = cs[J.seed];
            constexpr int OW = sizeof(ClimOct) / 16;
            if (cls < 2) {
                const int nslot = cls == 0 ? 4 : 14;
                for (int k = tid; k < nslot * OW; k += FT * FT) {
                    const int slot = proxy_slot(k / OW);
                    reinterpret_cast<uint4 *>(&C.oct[slot])[k % OW] = reinterpret_cast<const uint4 *>(&G.oct[slot])[k % OW];
                }
                constexpr int GW = sizeof(GradTab) / 16;
                if (tid < GW) reinterpret_cast<uint4 *>(&C.g)[tid] = reinterpret_cast<const uint4 *>(&G.g)[tid];
                if (tid < 6) C.factor[tid] = G.factor[tid];
            } else {
                const uint4 *src = reinterpret_cast<const uint4 *>(&G);
                uint4 *dst = reinterpret_cast<uint4 *>(&C);
                for (int k = tid; k < int(sizeof(ClimSeed) / 16); k += FT * FT) dst[k] = src[k];
            }
            __syncthreads();
            loaded_seed = J.seed;
            loaded_cls = cls;
        }
        // one cell per thread; wave w holds rows 2w (lanes 0-31) and 2w + 1 (lanes 32-63)
        const int nw = int(nwin(J));
        const int r = tid / FT, col = tid % FT;
        const int i = T.tx * FT + col, j = T.tz * FT + r;
        int v = 0;
        bool hot = false;
        if (i < nw && j < nw) {
            const int x = J.qx + (i - J.half) * J.step, z = J.qz + (j - J.half) * J.step;
            if (x >= -WORLD_Q && x <= WORLD_Q && z >= -WORLD_Q && z <= WORLD_Q) v = predicate(J.kind, C, g, x, z, J.cut, &hot);
        }
        const uint64_t mp = __ballot(v != 0), mc = __ballot(v == 2), mh = __ballot(hot);
        if ((tid & 63) == 0) {
            T.pass[r] = uint32_t(mp);
            T.pass[r + 1] = uint32_t(mp >> 32);
            T.cnt[r] = uint32_t(mc);
            T.cnt[r + 1] = uint32_t(mc >> 32);
            T.hot[r] = uint32_t(mh);
            T.hot[r + 1] = uint32_t(mh >> 32);
        }
    }
}

// Grow the reached set of tile ti to a fixpoint and hand the new frontier to the neighbours (one wave).
__device__ __forceinline__ void prop_tile(const FloodJob *jobs, FloodBufs &B, uint32_t ti, uint32_t *next_eval,
                                          uint32_t *next_prop, uint32_t *cnt_e, uint32_t *cnt_p) {
    Tile &T = B.tiles[ti];
    const int lane = threadIdx.x;
    const int r = lane & 31;
    if (lane == 0) atomicExch(&T.queued, 0u);
    __threadfence();
    __builtin_amdgcn_wave_barrier();
    const FloodJob J = jobs[T.job];
    const bool c8 = J.conn8 != 0;
    const uint32_t in = __atomic_load_n(&T.in[r], __ATOMIC_RELAXED);
    const uint32_t pass = T.pass[r], old = T.reach[r];
    uint32_t R = lane < FT ? (old | (in & pass)) : 0u;
    for (;;) {
        const uint32_t up = __shfl(R, (r + 31) & 31, 32), dn = __shfl(R, (r + 1) & 31, 32);
        const uint32_t upv = r > 0 ? up : 0u, dnv = r < FT - 1 ? dn : 0u;
        uint32_t D = R | (R << 1) | (R >> 1) | upv | dnv;
        if (c8) D |= (upv << 1) | (upv >> 1) | (dnv << 1) | (dnv >> 1);
        const uint32_t Rn = lane < FT ? (R | (D & pass)) : 0u;
        const bool ch = Rn != R;
        R = Rn;
        if (!__any(ch)) break;
    }
    const uint32_t NR = R & ~old;
    if (lane < FT && NR) T.reach[r] = R;
    if (!__any(NR != 0)) return;
    // ---- frontier across the tile edges
    const int nta = int(ntile_axis(J));
    const uint32_t up0 = __shfl(NR, 0, 32), dn31 = __shfl(NR, 31, 32);
    const uint32_t nr_m1 = __shfl(NR, (r + 31) & 31, 32), nr_p1 = __shfl(NR, (r + 1) & 31, 32);
    const uint32_t prev = r > 0 ? nr_m1 : 0u, next = r < FT - 1 ? nr_p1 : 0u;
    // left neighbour column 31 / right neighbour column 0, per row r
    bool lbit = (NR & 1u) != 0, rbit = (NR >> 31) != 0;
    if (c8) {
        lbit = lbit || (prev & 1u) || (next & 1u);
        rbit = rbit || (prev >> 31) || (next >> 31);
    }
    uint32_t vert_up = up0, vert_dn = dn31;
    if (c8) {
        vert_up |= (up0 << 1) | (up0 >> 1);
        vert_dn |= (dn31 << 1) | (dn31 >> 1);
    }
    // 8 directions: dx, dz, and this lane's contribution (row, mask); lane 0 collects the tiles to list
    uint32_t push_e[8], push_p[8];
    int ne = 0, np = 0;
    for (int d = 0; d < 8; d++) {
        const int dxs[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, dzs[8] = {0, 0, -1, 1, -1, -1, 1, 1};
        const int dx = dxs[d], dz = dzs[d];
        if (d >= 4 && !c8) break;
        int row = -1;
        uint32_t mask = 0;
        if (d == 0) { row = r; mask = lbit ? (1u << 31) : 0u; }
        else if (d == 1) { row = r; mask = rbit ? 1u : 0u; }
        else if (d == 2) { row = FT - 1; mask = lane == 0 ? vert_up : 0u; }
        else if (d == 3) { row = 0; mask = lane == 0 ? vert_dn : 0u; }
        else if (d == 4) { row = FT - 1; mask = (lane == 0 && (up0 & 1u)) ? (1u << 31) : 0u; }
        else if (d == 5) { row = FT - 1; mask = (lane == 0 && (up0 >> 31)) ? 1u : 0u; }
        else if (d == 6) { row = 0; mask = (lane == 0 && (dn31 & 1u)) ? (1u << 31) : 0u; }
        else { row = 0; mask = (lane == 0 && (dn31 >> 31)) ? 1u : 0u; }
        if (lane >= FT) mask = 0;
        const int ntx = T.tx + dx, ntz = T.tz + dz;
        if (ntx < 0 || ntz < 0 || ntx >= nta || ntz >= nta) continue;     // outside the window: edge (acc)
        mask &= col_mask(J, ntx);
        if (ntz * FT + row >= int(nwin(J))) mask = 0;
        if (!__any(mask != 0)) continue;
        uint32_t nti = 0;
        bool fresh = false;
        if (lane == 0) nti = tile_get(B, T.job, ntx, ntz, &fresh);
        nti = __shfl(nti, 0);
        fresh = __shfl(int(fresh), 0) != 0;
        if (nti == 0xffffffffu) {
            if (lane == 0) atomicOr(&B.lcount[FLOOD_OVF], 2u);
            continue;
        }
        Tile &N = B.tiles[nti];
        bool grew = false;
        if (mask) {
            const uint32_t o = atomicOr(&N.in[row], mask);
            grew = (o | mask) != o;
        }
        if (lane == 0 && fresh) push_e[ne++] = nti;
        if (__any(grew) && lane == 0) {
            __threadfence();
            if (atomicCAS(&N.queued, 0u, 1u) == 0u) push_p[np++] = nti;
        }
    }
    if (lane == 0) {
        if (ne) {
            const uint32_t k = atomicAdd(cnt_e, uint32_t(ne));
            for (int q = 0; q < ne; q++) {
                if (k + q < B.list_cap) next_eval[k + q] = push_e[q];
                else atomicOr(&B.lcount[FLOOD_OVF], 1u);
            }
        }
        if (np) {
            const uint32_t k = atomicAdd(cnt_p, uint32_t(np));
            for (int q = 0; q < np; q++) {
                if (k + q < B.list_cap) next_prop[k + q] = push_p[q];
                else atomicOr(&B.lcount[FLOOD_OVF], 1u);
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_flood_prop(const FloodJob *jobs, FloodBufs B, int round) {
    const int cur = round % 3, nxt = (round + 1) % 3;
    const uint32_t n = min(__atomic_load_n(&B.lcount[2 * cur + 1], __ATOMIC_RELAXED), B.list_cap);
    const uint32_t *list = B.lists[2 * cur + 1];
    uint32_t *next_eval = B.lists[2 * nxt], *next_prop = B.lists[2 * nxt + 1];
    uint32_t *cnt_e = &B.lcount[2 * nxt], *cnt_p = &B.lcount[2 * nxt + 1];
    for (uint32_t b = blockIdx.x; b < n; b += gridDim.x) prop_tile(jobs, B, list[b], next_eval, next_prop, cnt_e, cnt_p);
}

// Per-tile reduction of the flood outputs into res[job].
__global__ __launch_bounds__(64) void k_flood_acc(const FloodJob *jobs, FloodBufs B, uint32_t ntiles, FloodRes *res,
                                                  TileMask *masks, uint32_t *mcount, uint32_t mcap) {
    const uint32_t b = blockIdx.x;
    if (b >= ntiles) return;
    Tile &T = B.tiles[b];
    const FloodJob J = jobs[T.job];
    const int lane = threadIdx.x, r = lane & 31;
    const bool c8 = J.conn8 != 0;
    const uint32_t R = lane < FT ? T.reach[r] : 0u, I = lane < FT ? T.in[r] : 0u;
    const uint32_t up = __shfl(R, (r + 31) & 31, 32), dn = __shfl(R, (r + 1) & 31, 32);
    const uint32_t upv = r > 0 ? up : 0u, dnv = r < FT - 1 ? dn : 0u;
    uint32_t touched = R | I | (R << 1) | (R >> 1) | upv | dnv;
    if (c8) touched |= (upv << 1) | (upv >> 1) | (dnv << 1) | (dnv >> 1);
    const int nw = int(nwin(J));
    const int j = T.tz * FT + r;
    uint32_t win = (lane < FT && j < nw) ? col_mask(J, T.tx) : 0u;
    touched &= win;
    // world mask of this row
    const int z = J.qz + (j - J.half) * J.step;
    uint32_t world = 0;
    if (z >= -WORLD_Q && z <= WORLD_Q) {
        for (int c = 0; c < FT; c++) {
            const int x = J.qx + (T.tx * FT + c - J.half) * J.step;
            if (x >= -WORLD_Q && x <= WORLD_Q) world |= 1u << c;
        }
    }
    unsigned long long sampled = __popc(touched & world);
    unsigned long long hotc = __popc(touched & world & (lane < FT ? T.hot[r] : 0u));
    const bool border = (touched & ~world) != 0;
    const uint32_t RC = R & (lane < FT ? T.cnt[r] : 0u);
    unsigned long long cells = __popc(RC);
    int minx = INT_MAX, maxx = INT_MIN, minz = INT_MAX, maxz = INT_MIN;
    if (RC) {
        const int c0 = __ffs(RC) - 1, c1 = 31 - __clz(RC);
        minx = J.qx + (T.tx * FT + c0 - J.half) * J.step;
        maxx = J.qx + (T.tx * FT + c1 - J.half) * J.step;
        minz = maxz = z;
    }
    // a reached cell on the window boundary has a neighbour outside the window
    bool edge = false;
    if (R) {
        if (j == 0 || j == nw - 1) edge = true;
        if (T.tx == 0 && (R & 1u)) edge = true;
        const int last = nw - 1 - T.tx * FT;
        if (last >= 0 && last < FT && ((R >> last) & 1u)) edge = true;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sampled += __shfl_xor(sampled, o);
        hotc += __shfl_xor(hotc, o);
        cells += __shfl_xor(cells, o);
        minx = min(minx, __shfl_xor(minx, o));
        maxx = max(maxx, __shfl_xor(maxx, o));
        minz = min(minz, __shfl_xor(minz, o));
        maxz = max(maxz, __shfl_xor(maxz, o));
    }
    const bool any_border = __any(border), any_edge = __any(edge);
    if (lane == 0) {
        FloodRes &o = res[T.job];
        atomicAdd(&o.tiles, 1ull);
        if (sampled) atomicAdd(&o.sampled, sampled);
        if (hotc) atomicAdd(&o.hot, hotc);
        if (cells) {
            atomicAdd(&o.cells, cells);
            atomicMin(&o.minx, minx);
            atomicMax(&o.maxx, maxx);
            atomicMin(&o.minz, minz);
            atomicMax(&o.maxz, maxz);
        }
        if (any_border) atomicOr(&o.border, 1);
        if (any_edge) atomicOr(&o.edge, 1);
    }
    // counted cells of this tile -> mask export
    if (__any(RC != 0)) {
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(mcount, 1u);
        k = __shfl(k, 0);
        if (k < mcap) {
            if (lane < FT) masks[k].rows[lane] = RC;
            if (lane == 0) {
                masks[k].job = T.job;
                masks[k].tx = T.tx;
                masks[k].tz = T.tz;
            }
        }
    }
    // leave the tile and its hash slot clean for the next call (everything above has read them)
    __builtin_amdgcn_wave_barrier();
    if (lane < FT) {
        T.reach[lane] = 0;
        T.in[lane] = 0;
    }
    if (lane == 0) {
        if (T.slot1) B.hkeys[T.slot1 - 1] = HEMPTY;
        T.slot1 = 0;
        T.queued = 0;
    }
}

// ------------------------------------------------------------------ launchers

void launch_flood_resinit(hipStream_t st, const FloodJob *jobs, const uint32_t *jl, uint32_t n, FloodRes *res) {
    if (n) hipLaunchKernelGGL(k_flood_resinit, dim3((n + 255) / 256), dim3(256), 0, st, jobs, jl, n, res);
}
void launch_flood_start(hipStream_t st, const FloodJob *jobs, const uint32_t *jl, uint32_t n, FloodBufs B, uint32_t *eval,
                        uint32_t *prop) {
    if (n) hipLaunchKernelGGL(k_flood_start, dim3((n + 255) / 256), dim3(256), 0, st, jobs, jl, n, B, eval, prop);
}
void launch_flood_round(hipStream_t st, const FloodJob *jobs, const ClimSeed *cs, const SurfGrid &g, FloodBufs B, int round,
                        uint32_t width, hipEvent_t ev0, hipEvent_t ev1) {
    const uint32_t w = width < 8 ? 8 : width;
    const dim3 eg(w < EVAL_GRID ? w : EVAL_GRID);
    if (ev0) hipExtLaunchKernelGGL(k_flood_eval, eg, dim3(FT * FT), 0, st, ev0, ev1, 0u, jobs, cs, g, B, round);
    else hipLaunchKernelGGL(k_flood_eval, eg, dim3(FT * FT), 0, st, jobs, cs, g, B, round);
    hipLaunchKernelGGL(k_flood_prop, dim3(w < PROP_GRID ? w : PROP_GRID), dim3(64), 0, st, jobs, B, round);
}
void launch_flood_acc(hipStream_t st, const FloodJob *jobs, FloodBufs B, uint32_t ntiles, FloodRes *res, TileMask *masks,
                      uint32_t *mcount, uint32_t mcap) {
    if (ntiles) hipLaunchKernelGGL(k_flood_acc, dim3(ntiles), dim3(64), 0, st, jobs, B, ntiles, res, masks, mcount, mcap);
}

}  // namespace dsx
