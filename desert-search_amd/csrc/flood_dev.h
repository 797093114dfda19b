// Device helpers of the tiled flood engine shared by its kernels (floods.hip) and the small-window
// engine's hand-over (flood_lds.hip): window geometry, agent-scope accesses, the tile hash, the work
// queue and tile ownership. The protocol is described at the top of floods.hip.
#pragma once
#include "floods.h"

namespace dsx {

__device__ __forceinline__ uint32_t nwin(const FloodJob &J) { return uint32_t(2 * J.half + 1); }
__device__ __forceinline__ uint32_t ntile_axis(const FloodJob &J) { return (nwin(J) + FT - 1) / FT; }
// valid-column mask of a tile row: columns whose window index is < n
__device__ __forceinline__ uint32_t col_mask(const FloodJob &J, int tx) {
    const int n = int(nwin(J)), c0 = tx * FT;
    const int k = n - c0;
    return k >= FT ? 0xffffffffu : (k <= 0 ? 0u : ((1u << k) - 1u));
}

// agent-scope (all XCDs) relaxed loads and stores of tile state shared inside a kernel
template <class T> __device__ __forceinline__ T ald(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void ast(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Hand-over points: every earlier memory operation of this wave has completed (its agent-scope
// stores and atomics are performed at the coherence point) before the next one issues. Shared state
// is only ever touched through agent-scope accesses, so no cache writeback or invalidation is needed
// (an agent-scope release fence would write back the whole L2 at every hand-over).
__device__ __forceinline__ void mem_done() { __builtin_amdgcn_s_waitcnt(0); }

// Words that kernels update with atomic read-modify-writes (control counters, tile queue states and
// `in` masks, hash keys, flood results) are also set and reset with atomics, never with stores: every
// write to such a word then takes the same path to memory, so a reset cannot land after a later update.
__device__ __forceinline__ void aset(uint32_t *p, uint32_t v) { (void)atomicExch(p, v); }
__device__ __forceinline__ void aset(int *p, int v) { (void)atomicExch(p, v); }
__device__ __forceinline__ void aset(unsigned long long *p, unsigned long long v) { (void)atomicExch(p, v); }
// Flood results (acc adds to them with atomics) are initialised word by word with atomic exchanges.
__device__ __forceinline__ void res_store(FloodRes *p, const FloodRes &r) {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(&r);
    uint32_t *d = reinterpret_cast<uint32_t *>(p);
#pragma unroll
    for (int i = 0; i < int(sizeof(FloodRes) / 4); i++) aset(&d[i], s[i]);
}
__device__ __forceinline__ void mask_head_store(TileMask &m, uint32_t job, int32_t tx, int32_t tz) {
    ast(&m.job, job);
    ast(&m.tx, tx);
    ast(&m.tz, tz);
}
// Exported mask k of a batch, written by one wave (lane < 32: its row): into the device list (k < mcap) and,
// when the batch asks for it, into pinned host memory as well (k < B.h_mcap).
__device__ __forceinline__ void mask_export(const FloodBufs &B, TileMask *masks, uint32_t mcap, uint32_t k, int lane,
                                            uint32_t row, uint32_t job, int32_t tx, int32_t tz) {
    if (k < mcap) {
        if (lane < 32) ast(&masks[k].rows[lane], row);
        if (lane == 0) mask_head_store(masks[k], job, tx, tz);
    }
    if (B.h_masks && k < B.h_mcap) {
        TileMask &h = B.h_masks[k];
        if (lane < 32) h.rows[lane] = row;
        if (lane == 0) {
            h.job = job;
            h.tx = tx;
            h.tz = tz;
        }
    }
}

constexpr unsigned long long HEMPTY = ~0ull;
__device__ __forceinline__ unsigned long long hkey(uint32_t job, int tx, int tz) {
    return (uint64_t(job) << 44) | (uint64_t(uint32_t(tx) & 2047u) << 33) | (uint64_t(uint32_t(tz) & 2047u) << 22);
}
// Returns the tile index of (job, tx, tz), allocating it if absent; *fresh = this call allocated it.
// Keys and values are published together in one 64-bit CAS, so no reader ever sees a half-made entry.
__device__ inline uint32_t tile_get(const FloodBufs &B, uint32_t job, int tx, int tz, bool *fresh) {
    const unsigned long long key = hkey(job, tx, tz);
    uint32_t h = uint32_t(mix64(key) & B.hmask);
    uint32_t mine = 0xffffffffu;
    *fresh = false;
    for (uint32_t probe = 0; probe <= B.hmask; probe++, h = (h + 1) & B.hmask) {
        unsigned long long cur = ald(&B.hkeys[h]);
        if (cur == HEMPTY) {
            if (mine == 0xffffffffu) {
                mine = atomicAdd(B.ntiles, 1u);
                if (mine >= B.tile_cap) return 0xffffffffu;        // pool exhausted (reported as OVF_TILES)
                Tile &t = B.tiles[mine];
                ast(&t.job, job);
                ast(&t.tx, tx);
                ast(&t.tz, tz);
                mem_done();
            }
            const unsigned long long want = key | mine;
            cur = atomicCAS(&B.hkeys[h], HEMPTY, want);
            if (cur == HEMPTY) {
                ast(&B.tiles[mine].slot1, h + 1);
                *fresh = true;
                return mine;
            }
        }
        if ((cur & ~0x3fffffull) == key) return uint32_t(cur & 0x3fffffull);
    }
    return 0xffffffffu;
}

// Append tile ti to the work queue (counted as pending first).
__device__ __forceinline__ void wq_push(const FloodBufs &B, uint32_t ti) {
    atomicAdd(&B.ctl[Q_PENDING], 1u);
    const uint32_t k = atomicAdd(&B.ctl[Q_TAIL], 1u);
    if (k >= B.wq_cap) {
        atomicOr(&B.ctl[FLOOD_OVF], OVF_QUEUE);
        return;
    }
    mem_done();
    __hip_atomic_store(&B.wq[k], ti, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// After new `in` bits of tile ti are published: make sure the tile is (re)processed.
__device__ __forceinline__ void tile_touch(const FloodBufs &B, uint32_t ti) {
    uint32_t *q = &B.tiles[ti].qstate;
    mem_done();
    for (;;) {
        uint32_t old = atomicCAS(q, 0u, 1u);
        if (old == 0u) {
            wq_push(B, ti);
            return;
        }
        if (old == 2u) return;
        old = atomicCAS(q, 1u, 2u);   // queued or in process: have it run again
        if (old != 0u) return;        // now 2; if it went idle meanwhile (0), queue it ourselves
    }
}

// Touch the start cell (half, half) of flood `job` and queue its tile.
__device__ inline void flood_start_tile(const FloodJob &J, uint32_t job, const FloodBufs &B) {
    const int c = J.half % FT, t = J.half / FT;
    bool fresh = false;
    const uint32_t ti = tile_get(B, job, t, t, &fresh);
    if (ti == 0xffffffffu) {
        atomicOr(&B.ctl[FLOOD_OVF], OVF_TILES);
        return;
    }
    atomicOr(&B.tiles[ti].in[c], 1u << c);
    tile_touch(B, ti);
}

}  // namespace dsx
