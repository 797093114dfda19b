// Tiled flood engine for gfx950 (bounded bridged floods, dense_flood.hpp:52-86), one persistent
// kernel per batch of floods.
//
// The reference floods depth-first through a dense (2*half+1)^2 byte window. Every output of that
// flood (cells, sampled, bbox, edge, border, and the final window state) depends only on the set of
// cells reachable from the start through predicate != 0 cells and on their neighbourhood, not on the
// visiting order. This engine computes the same sets tile-parallel:
//   * the window is split into 32x32 tiles, allocated on first touch through a device hash table;
//   * k_flood_run is a persistent kernel over a device work queue of tiles. A workgroup takes a tile
//     and grows its reached set to a fixpoint with bit-parallel dilation (one wave, one lane per 32-bit
//     row mask) through the cells evaluated so far; whenever the touched cells (the reached set's
//     neighbours and the cells touched from neighbouring tiles) reach an 8x8 sub-block not evaluated
//     yet, its waves evaluate the FP64 predicate of those sub-blocks (one cell per lane, the octaves
//     the predicate reads staged in LDS) and the fixpoint continues. Then it ORs the new frontier
//     across the tile edges into the neighbours' `in` masks and queues the neighbours it touched. It
//     runs until no work is pending: no per-round kernel launches and no host round trips;
//   * k_flood_acc then reduces each tile:
//       reached = pushed cells, touched = reached | in | dilate_inside(reached),
//       cells = |reached & pred==2|, sampled = |touched inside the world|, border = touched outside
//       the world, edge = reached cell on the window boundary, bbox over counted reached cells
//       (starting at the window centre, as dense_flood.hpp:57).
//
// Ownership and visibility. A tile is processed by one workgroup at a time: its qstate word goes
// 0 -> 1 when it is queued, and a workgroup that touches a tile in state 1 (queued or in process)
// moves it to 2, which makes the processing workgroup run it again before it returns it to 0. Tile
// state shared between workgroups is only read and written with agent-scope accesses (sc1) and
// atomics, which are coherent across the XCDs' L2 caches; at every hand-over (queue slot, qstate,
// hash publication) the wave first waits until its earlier accesses have completed.
//
// Termination. Q_PENDING counts queued items not yet finished (an item's own pushes are counted
// before it finishes), so it reaches 0 exactly when no more work can appear. A workgroup claims the
// next queue slot and polls it (each waiter on its own address); it leaves when Q_PENDING is 0, when
// an overflow was flagged, or after a time limit of 5 s (flagged as an error; the host then runs the batch
// once more on a reset engine), so every wave exits.
#include "floods.h"
#include "flood_dev.h"
#include <hip/hip_ext.h>
#include <type_traits>
#include <algorithm>
#include <climits>
#include <cstdlib>

namespace dsx {

// Jobs are written and read word by word with agent-scope accesses where the same kernel both stages
// and floods (a plain store is not visible to workgroups on other XCDs before the kernel ends).
__device__ __forceinline__ void job_store(FloodJob *p, const FloodJob &J) {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(&J);
    uint32_t *d = reinterpret_cast<uint32_t *>(p);
#pragma unroll
    for (int i = 0; i < int(sizeof(FloodJob) / 4); i++) ast(&d[i], s[i]);
}
__device__ __forceinline__ FloodJob job_load(const FloodJob *p) {
    FloodJob J;
    uint32_t *d = reinterpret_cast<uint32_t *>(&J);
    const uint32_t *s = reinterpret_cast<const uint32_t *>(p);
#pragma unroll
    for (int i = 0; i < int(sizeof(FloodJob) / 4); i++) d[i] = ald(&s[i]);
    return J;
}

// Item q of a batch's staging, in two parts: stage_job copies job q from the pinned staging to the device
// and initialises its result (q < n) and copies small-window list entry q (q < n_small); stage_tiles touches
// and queues tiled flood q's start tile (q < n_big) and applies hint q (q < nh). A queued tile's job must be
// on the device before any workgroup can take the tile: where the same kernel stages and floods, every job
// is stored before any tile is queued.
__device__ __forceinline__ void stage_job(uint32_t q, const uint8_t *h_up, uint32_t n, uint32_t n_small, uint8_t *d_up,
                                          FloodRes *res) {
    const FloodJob *hj = reinterpret_cast<const FloodJob *>(h_up);
    FloodJob *dj = reinterpret_cast<FloodJob *>(d_up);
    const uint32_t *hl = reinterpret_cast<const uint32_t *>(hj + n);
    uint32_t *dl = reinterpret_cast<uint32_t *>(dj + n);
    if (q < n) {
        const FloodJob J = hj[q];
        job_store(&dj[q], J);
        FloodRes r{};
        r.minx = r.maxx = J.qx;
        r.minz = r.maxz = J.qz;
        res_store(&res[q], r);
        if (q < n_small) dl[q] = hl[q];
    }
}
__device__ __forceinline__ void stage_tiles(uint32_t q, const uint8_t *h_up, uint32_t n, uint32_t n_big, uint32_t nh,
                                            uint8_t *d_up, const FloodBufs &B) {
    const FloodJob *hj = reinterpret_cast<const FloodJob *>(h_up);
    const uint32_t *hl = reinterpret_cast<const uint32_t *>(hj + n);
    uint32_t *dl = reinterpret_cast<uint32_t *>(reinterpret_cast<FloodJob *>(d_up) + n);
    if (q < n_big) {
        const uint32_t k = hl[n + q];
        dl[n + q] = k;
        flood_start_tile(hj[k], k, B);
    }
    if (q < nh) {   // hinted tiles: allocated, hinted and queued; their jobs are tiled jobs of this batch
        const FloodHint hnt = reinterpret_cast<const FloodHint *>(hl + 2 * n)[q];
        const FloodJob J = hj[hnt.job];
        const int nta = int(ntile_axis(J));
        if (hnt.blocks && hnt.tx >= 0 && hnt.tz >= 0 && hnt.tx < nta && hnt.tz < nta) {
            bool fresh = false;
            const uint32_t ti = tile_get(B, hnt.job, hnt.tx, hnt.tz, &fresh);
            if (ti == 0xffffffffu) {
                atomicOr(&B.ctl[FLOOD_OVF], OVF_TILES);
            } else {
                atomicOr(&B.tiles[ti].hint, hnt.blocks & 0xffffu);
                tile_touch(B, ti);
            }
        }
    }
}

// (the tiled engine that consumes the queued tiles is a later kernel, so its jobs are all in place)
__global__ __launch_bounds__(256) void k_flood_stage(const uint8_t *h_up, uint32_t n, uint32_t n_small, uint32_t n_big,
                                                     uint32_t nh, uint8_t *d_up, FloodRes *res, FloodBufs B) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    stage_job(q, h_up, n, n_small, d_up, res);
    stage_tiles(q, h_up, n, n_big, nh, d_up, B);
    if (q == 0) aset(&B.ctl[STAGE_SEQ], B.seq + 1);
}

// Sub-blocks: a tile's 32x32 cells are evaluated lazily in 16 blocks of 8x8 (bit bz*4 + bx of the
// tile's `evald` mask), only those holding cells the flood touches.
constexpr int SB = 8, NSB = FT / SB;

// Evaluate sub-block sb of the tile at (tx, tz) with one wave (one cell per lane) and OR its bits into
// the LDS row masks (pass = v != 0, cnt = v == 2, hot = the predicate's temperature test passed; cells
// outside the window or the world are none of them).
template <int PRE, int LIGHT, class CS>
__device__ __forceinline__ void eval_subblock(int sb, int tx, int tz, const FloodJob &J, const CS &C,
                                              const SurfGrid &g, uint32_t *s_pass, uint32_t *s_cnt, uint32_t *s_hot) {
    const int lane = threadIdx.x & 63;
    const int bx = sb % NSB, bz = sb / NSB;
    const int r = bz * SB + (lane >> 3), col = bx * SB + (lane & 7);
    const int i = tx * FT + col, j = tz * FT + r;
    const int nw = int(nwin(J));
    int v = 0;
    bool hot = false;
    if (i < nw && j < nw) {
        const int x = J.qx + (i - J.half) * J.step, z = J.qz + (j - J.half) * J.step;
        if (x >= -WORLD_Q && x <= WORLD_Q && z >= -WORLD_Q && z <= WORLD_Q)
        {
            if constexpr (LIGHT)   // hot-zone and proxy floods only: no biome lookup in this instance
                v = J.kind == PRED_HOT ? (PRE ? hot_state_pre(C, x, z) : hot_state(C, x, z))
                                       : (PRE ? proxy_state_pre(C, x, z, J.cut, &hot) : proxy_state(C, x, z, J.cut, &hot));
            else
                v = PRE ? predicate_pre(J.kind, C, g, x, z, J.cut, &hot) : predicate(J.kind, C, g, x, z, J.cut, &hot);
        }
    }
    const uint64_t mp = __ballot(v != 0), mc = __ballot(v == 2), mh = __ballot(hot);
    if (lane < SB) {   // lane rr: row bz*8 + rr of the sub-block
        const int sh = lane * SB;
        const int row = bz * SB + lane;
        const uint32_t p = uint32_t((mp >> sh) & 0xffu) << (bx * SB), c = uint32_t((mc >> sh) & 0xffu) << (bx * SB),
                       h = uint32_t((mh >> sh) & 0xffu) << (bx * SB);
        if (p) atomicOr(&s_pass[row], p);
        if (c) atomicOr(&s_cnt[row], c);
        if (h) atomicOr(&s_hot[row], h);
    }
}

// The light instance's FP64 evaluation by pairs of waves: the pair's first wave takes side 0 of every noise,
// its second side 1, for the same 64 cells (one per lane), which halves the octave samples per lane (T: 2 of
// 4; a hot proxy cell's C and E: 5 of 10). Side 1's sums and terms reach side 0 through X, and side 0
// finishes each value in noise_sample's and noise_octaves' operation order (so the states are those of
// hot_state / proxy_state). sb < 0: the pair has no sub-block this pass but takes part in the block's
// barriers, which every wave of the block must reach the same number of times.
struct PairX {
    double t1[64];       // side 1's T sum
    double c1[3][64];    // side 1's weighted C terms (octaves 0-2)
    double e1[2][64];    // side 1's weighted E terms (octaves 0-1)
    uint32_t hot[64];    // side 0's temperature test, for side 1
};
template <class CS>
__device__ __forceinline__ void eval_pair(int sb, int side, int tx, int tz, const FloodJob &J, const CS &C, PairX &X,
                                          uint32_t *s_pass, uint32_t *s_cnt, uint32_t *s_hot) {
    const int lane = threadIdx.x & 63;
    const int bx = sb < 0 ? 0 : sb % NSB, bz = sb < 0 ? 0 : sb / NSB;
    const int r = bz * SB + (lane >> 3), col = bx * SB + (lane & 7);
    const int i = tx * FT + col, j = tz * FT + r;
    const int nw = int(nwin(J));
    int x = 0, z = 0;
    bool in = false;
    if (sb >= 0 && i < nw && j < nw) {
        x = J.qx + (i - J.half) * J.step;
        z = J.qz + (j - J.half) * J.step;
        in = x >= -WORLD_Q && x <= WORLD_Q && z >= -WORLD_Q && z <= WORLD_Q;
    }
    // T: each side's sum, side 1's to side 0
    const double ts = in ? noise_side(C, DSX_CH_T, side, x, z) : 0.0;
    if (side == 1) X.t1[lane] = ts;
    __syncthreads();
    bool hot = false;
    int v = 0;
    if (side == 0 && in) {
        const double T = (ts + X.t1[lane]) * C.factor[DSX_CH_T];
        hot = T >= 0.55;
        if (J.kind == PRED_HOT) v = hot ? 2 : 0;
    }
    if (J.kind == PRED_PROXY) {   // (uniform: one tile's job)
        if (side == 0) X.hot[lane] = hot ? 1u : 0u;
        __syncthreads();
        const bool h = side == 0 ? hot : X.hot[lane] != 0u;
        double cs = 0, es = 0;   // side 0: C's and E's running sums over its octaves
        if (in && h) {
            if (side == 0) {
                cs = octaves_side(C, DSX_CH_C, 0, x, z, 3, nullptr);
                es = octaves_side(C, DSX_CH_E, 0, x, z, 2, nullptr);
            } else {
                double tc[3], te[2];
                octaves_side(C, DSX_CH_C, 1, x, z, 3, tc);
                octaves_side(C, DSX_CH_E, 1, x, z, 2, te);
#pragma unroll
                for (int k = 0; k < 3; k++) X.c1[k][lane] = tc[k];
#pragma unroll
                for (int k = 0; k < 2; k++) X.e1[k][lane] = te[k];
            }
        }
        __syncthreads();
        if (side == 0 && in) {
            float prob = 0.f;
            if (hot) {
                double c = cs, e = es;   // noise_octaves: side 0's sum, side 1's terms in order, times factor
                for (int k = 0; k < 3; k++) c += X.c1[k][lane];
                for (int k = 0; k < 2; k++) e += X.e1[k][lane];
                c = c * C.factor[DSX_CH_C];
                e = e * C.factor[DSX_CH_E];
                int ii = int(floor((c + 0.8) / 0.05)), jj = int(floor((e + 1.0) / 0.05));
                ii = ii < 0 ? 0 : (ii > 39 ? 39 : ii);
                jj = jj < 0 ? 0 : (jj > 39 ? 39 : jj);
                prob = DSX_DPROB[ii][jj];
            }
            v = prob >= J.cut ? 2 : 0;
        }
    }
    __syncthreads();   // X is free for the pair's next sub-block
    if (side == 0 && sb >= 0) {
        const uint64_t mp = __ballot(v != 0), mc = __ballot(v == 2), mh = __ballot(J.kind == PRED_PROXY && hot);
        if (lane < SB) {   // lane rr: row bz*8 + rr of the sub-block
            const int sh = lane * SB;
            const int row = bz * SB + lane;
            const uint32_t p = uint32_t((mp >> sh) & 0xffu) << (bx * SB), c = uint32_t((mc >> sh) & 0xffu) << (bx * SB),
                           h = uint32_t((mh >> sh) & 0xffu) << (bx * SB);
            if (p) atomicOr(&s_pass[row], p);
            if (c) atomicOr(&s_cnt[row], c);
            if (h) atomicOr(&s_hot[row], h);
        }
    }
}

// One wave: grow the reached set of tile T through its evaluated passable cells to a fixpoint. Returns
// the sub-blocks that hold touched cells not yet evaluated (reached cells' neighbours, and cells
// touched from neighbouring tiles). When none are left (0), the reached set is final: it is stored
// and its new frontier is handed to the neighbouring tiles.
__device__ uint32_t prop_tile(const FloodBufs &B, uint32_t ti, const FloodJob &J, const uint32_t *s_pass, uint32_t evald,
                              uint32_t hint) {
    Tile &T = B.tiles[ti];
    const int lane = threadIdx.x & 63;
    const int r = lane & 31;
    const bool c8 = (J.conn8 & 1) != 0;
    const uint32_t job = ald(&T.job);
    const int tx = ald(&T.tx), tz = ald(&T.tz);
    const uint32_t in = lane < FT ? ald(&T.in[r]) : 0u;
    const uint32_t pass = lane < FT ? s_pass[r] : 0u, old = lane < FT ? ald(&T.reach[r]) : 0u;
    uint32_t R = lane < FT ? (old | (in & pass)) : 0u;
    uint32_t D;
    for (;;) {
        const uint32_t up = __shfl(R, (r + 31) & 31, 32), dn = __shfl(R, (r + 1) & 31, 32);
        const uint32_t upv = r > 0 ? up : 0u, dnv = r < FT - 1 ? dn : 0u;
        D = R | (R << 1) | (R >> 1) | upv | dnv;
        if (c8) D |= (upv << 1) | (upv >> 1) | (dnv << 1) | (dnv >> 1);
        const uint32_t Rn = lane < FT ? (R | (D & pass)) : 0u;
        const bool ch = Rn != R;
        R = Rn;
        if (!__any(ch)) break;
    }
    // touched cells of this tile (inside the window) and the sub-blocks they need
    const int nw = int(nwin(J));
    const uint32_t win = (lane < FT && tz * FT + r < nw) ? col_mask(J, tx) : 0u;
    const uint32_t touched = (D | in) & win;
    uint32_t need = 0;
#pragma unroll
    for (int bx = 0; bx < NSB; bx++) {
        const uint64_t rows = __ballot(((touched >> (bx * SB)) & 0xffu) != 0u);
#pragma unroll
        for (int bz = 0; bz < NSB; bz++)
            if ((rows >> (bz * SB)) & 0xffu) need |= 1u << (bz * NSB + bx);
    }
    need = (need | hint) & ~evald;
    if (need) return need;
    const uint32_t NR = R & ~old;
    if (lane < FT && NR) ast(&T.reach[r], R);
    if (!__any(NR != 0)) return 0u;
    // ---- frontier across the tile edges, all directions at once: the wave's lanes look up the
    // neighbour tiles in parallel (lane d: direction d), OR the crossing rows into their `in` masks in
    // two atomic instructions (lanes 0-31: the left neighbour's rows, 32-63: the right one's; lanes 0-5:
    // the rows above, below and diagonal), then queue the neighbours that grew, again lane d for d.
    // Directions d: 0 left, 1 right, 2 up, 3 down, 4 up-left, 5 up-right, 6 down-left, 7 down-right.
    const int nta = int(ntile_axis(J));
    const int rowv = lane & 31;
    // (every lane takes part in each shuffle: a lane reading an inactive lane's value gets nothing)
    const uint32_t nr = __shfl(NR, rowv), nr_m1 = __shfl(NR, (rowv + FT - 1) & (FT - 1)),
                   nr_p1 = __shfl(NR, (rowv + 1) & (FT - 1));
    const uint32_t prv = rowv > 0 ? nr_m1 : 0u, nxt = rowv < FT - 1 ? nr_p1 : 0u;
    const uint32_t up0 = __shfl(NR, 0), dn31 = __shfl(NR, FT - 1);
    auto dxd = [](int d) { return d == 0 || d == 4 || d == 6 ? -1 : (d == 1 || d == 5 || d == 7 ? 1 : 0); };
    auto dzd = [](int d) { return d == 2 || d == 4 || d == 5 ? -1 : (d == 3 || d == 6 || d == 7 ? 1 : 0); };
    // the mask this lane ORs into row `row` of the neighbour in direction d (0 if none)
    auto clip = [&](int d, int row, uint32_t m) -> uint32_t {
        const int ntx = tx + dxd(d), ntz = tz + dzd(d);
        if (ntx < 0 || ntz < 0 || ntx >= nta || ntz >= nta) return 0u;   // outside the window: edge (acc)
        if (ntz * FT + row >= nw) return 0u;
        return m & col_mask(J, ntx);
    };
    const int sd = lane < 32 ? 0 : 1;   // this lane's side direction, row rowv
    uint32_t smask;
    if (sd == 0) smask = ((nr & 1u) || (c8 && ((prv & 1u) || (nxt & 1u)))) ? (1u << 31) : 0u;
    else smask = ((nr >> 31) || (c8 && ((prv >> 31) || (nxt >> 31)))) ? 1u : 0u;
    smask = clip(sd, rowv, smask);
    const int vd = lane + 2;            // lanes 0-5: direction lane + 2
    int vrow = 0;
    uint32_t vmask = 0;
    if (lane < 6 && (lane < 2 || c8)) {
        uint32_t vu = up0, vdn = dn31;
        if (c8) {
            vu |= (up0 << 1) | (up0 >> 1);
            vdn |= (dn31 << 1) | (dn31 >> 1);
        }
        if (vd == 2) { vrow = FT - 1; vmask = vu; }
        else if (vd == 3) { vrow = 0; vmask = vdn; }
        else if (vd == 4) { vrow = FT - 1; vmask = (up0 & 1u) ? (1u << 31) : 0u; }
        else if (vd == 5) { vrow = FT - 1; vmask = (up0 >> 31) ? 1u : 0u; }
        else if (vd == 6) { vrow = 0; vmask = (dn31 & 1u) ? (1u << 31) : 0u; }
        else { vrow = 0; vmask = (dn31 >> 31) ? 1u : 0u; }
        vmask = clip(vd, vrow, vmask);
    }
    const uint64_t bs = __ballot(smask != 0u), bv = __ballot(vmask != 0u);
    uint32_t want = (uint32_t(bs) ? 1u : 0u) | (uint32_t(bs >> 32) ? 2u : 0u) | (uint32_t(bv & 0x3fu) << 2);
    if (!want) return 0u;
    // neighbour tiles, looked up (or allocated) by lane d in parallel
    uint32_t nti = 0xffffffffu;
    bool fresh = false;
    if (lane < 8 && ((want >> lane) & 1u)) {
        nti = tile_get(B, job, tx + dxd(lane), tz + dzd(lane), &fresh);
        if (nti == 0xffffffffu) atomicOr(&B.ctl[FLOOD_OVF], OVF_TILES);
    }
    want &= ~uint32_t(__ballot(lane < 8 && nti == 0xffffffffu));
    const uint32_t nti_s = __shfl(nti, sd), nti_v = __shfl(nti, lane < 6 ? vd : 0);
    bool gs = false, gv = false;
    if (smask && ((want >> sd) & 1u)) {
        const uint32_t o = atomicOr(&B.tiles[nti_s].in[rowv], smask);
        gs = (o | smask) != o;
    }
    if (vmask && ((want >> vd) & 1u)) {
        const uint32_t o = atomicOr(&B.tiles[nti_v].in[vrow], vmask);
        gv = (o | vmask) != o;
    }
    const uint64_t gsb = __ballot(gs), gvb = __ballot(gv);
    const uint32_t grew = (uint32_t(gsb) ? 1u : 0u) | (uint32_t(gsb >> 32) ? 2u : 0u) | (uint32_t(gvb & 0x3fu) << 2);
    if (lane < 8 && ((grew >> lane) & 1u)) tile_touch(B, nti);
    return 0u;
}

// Persistent workgroups over the work queue (NT threads each; wave 0 propagates, all waves evaluate).
// OCC: waves per SIMD the register allocation must allow. LIGHT: an instance for batches whose tiled
// floods are all hot-zone or proxy floods; it leaves out the biome predicates and so needs fewer
// registers (more workgroups resident beside the other streams' kernels).
// With sa.nst > 0 the kernel stages its batch first: the first sa.nst workgroups to arrive (by a ticket
// drawn on arrival, so each of them is running) store every job and result, wait for each other, queue
// the start and hinted tiles (grid-stride loops) and count themselves done; no workgroup leaves before all
// of them have, since the items they queue are counted as pending before that.
template <int NT, int OCC, int PRE, int LIGHT>
__global__ __launch_bounds__(NT, OCC) void k_flood_run(const FloodJob *jobs, const ClimSeed *cs, SurfGrid g, FloodBufs B,
                                                       int backoff, int spec, int tprof, StageArgs sa) {
    constexpr int NWAVE = NT / 64;
    // spec bit 1: raise this kernel's waves' issue priority (s_setprio) above the default of other
    // kernels sharing the SIMDs, e.g. the VALU-bound gate; bit 0: speculative sub-blocks
    if (spec & 2) __builtin_amdgcn_s_setprio(2);
    const int spec_pairs = spec & 4;
    spec &= 1;
    const uint64_t t_born = tprof ? __builtin_amdgcn_s_memrealtime() : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // what the first workgroup finds (time-limit diagnostics)
        aset(&B.ctl[RUN_SEQ], B.seq + 1);
        aset(&B.ctl[RUN_TAIL0], atomicAdd(&B.ctl[Q_TAIL], 0u));
        aset(&B.ctl[RUN_STAGE_SEQ], atomicAdd(&B.ctl[STAGE_SEQ], 0u));
    }
    // the octaves the predicates read: the proxy's 14 (ClimP, 5 KB of LDS) in the light instance
    __shared__ std::conditional_t<LIGHT != 0, ClimP, ClimSeed> C;
    __shared__ uint32_t s_pass[FT], s_cnt[FT], s_hot[FT];
    __shared__ uint32_t s_item, s_again, s_need, s_evald, s_hint;
    __shared__ int s_stop;
    // (light instance, FP64) sub-blocks evaluated by pairs of waves, one noise side each (eval_pair);
    // DSX_FLOOD_PAIR=0 keeps one wave per sub-block (spec bit 2 off)
    constexpr bool PAIRS = LIGHT && PRE == 0 && NWAVE >= 2 && NWAVE % 2 == 0;
    __shared__ PairX s_x[PAIRS ? NWAVE / 2 : 1];
    const bool pairs = PAIRS && spec_pairs != 0;
    const int tid = threadIdx.x, wave = tid >> 6;
    int loaded_seed = -1, loaded_cls = -1;
    if (sa.nst) {
        if (tid == 0) s_item = atomicAdd(&B.ctl[STAGE_TICKET], 1u);
        __syncthreads();
        const uint32_t t = s_item;
        if (t < sa.nst) {
            // every job and result first, in all staging workgroups (each of them is running, so this wait
            // ends), then the start and hinted tiles: no tile is queued before its job is on the device
            for (uint32_t q = t * NT + tid; q < sa.n; q += sa.nst * NT) stage_job(q, sa.h_up, sa.n, 0u, sa.d_up, sa.res);
            mem_done();
            __syncthreads();
            if (tid == 0) {
                atomicAdd(&B.ctl[STAGE_JOBS], 1u);
                const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
                while (ald(&B.ctl[STAGE_JOBS]) < sa.nst) {
                    if (__builtin_amdgcn_s_memrealtime() - t0 > 500000000ull) {   // 5 s at 100 MHz: flagged
                        atomicOr(&B.ctl[FLOOD_OVF], OVF_TIME);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __syncthreads();
            const uint32_t m = max(sa.n_big, sa.nh);
            for (uint32_t q = t * NT + tid; q < m; q += sa.nst * NT) stage_tiles(q, sa.h_up, sa.n, sa.n_big, sa.nh, sa.d_up, B);
            mem_done();   // this wave's queue pushes and pending counts have completed
            __syncthreads();
            if (tid == 0) {
                if (t == 0) aset(&B.ctl[STAGE_SEQ], B.seq + 1);
                atomicAdd(&B.ctl[STAGE_DONE], 1u);
            }
        }
        __syncthreads();   // s_item is taken again below
    }
    for (;;) {
        // ---- take the next queue slot and wait for its tile
        if (tid == 0) {
            uint32_t item = Q_EMPTY;
            int stop = 0;
            const uint32_t k = atomicAdd(&B.ctl[Q_HEAD], 1u);
            const uint64_t tw0 = tprof ? __builtin_amdgcn_s_memrealtime() : 0;
            if (k >= B.wq_cap) {
                stop = 1;
            } else {
                const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
                for (uint32_t it = 0;; it++) {
                    item = __hip_atomic_load(&B.wq[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    if (item != Q_EMPTY) break;
                    if ((it & 7u) == 7u) {   // the exit conditions, every 8th poll
                        // staging complete (read first: its queued items are counted as pending by then),
                        // then nothing pending
                        bool staged = true;
                        if (sa.nst) {
                            staged = ald(&B.ctl[STAGE_DONE]) == sa.nst;
                            mem_done();
                        }
                        if ((staged && ald(&B.ctl[Q_PENDING]) == 0u) || ald(&B.ctl[FLOOD_OVF]) != 0u) {
                            stop = 1;
                            break;
                        }
                        if (__builtin_amdgcn_s_memrealtime() - t0 > 500000000ull) {   // 5 s at 100 MHz
                            if (atomicCAS(&B.ctl[TIME_SEEN], 0u, 1u) == 0u) {
                                aset(&B.ctl[TIME_SLOT], k);
                                aset(&B.ctl[TIME_HEAD], atomicAdd(&B.ctl[Q_HEAD], 0u));
                                aset(&B.ctl[TIME_TAIL], atomicAdd(&B.ctl[Q_TAIL], 0u));
                                aset(&B.ctl[TIME_PENDING], atomicAdd(&B.ctl[Q_PENDING], 0u));
                                aset(&B.ctl[TIME_ITEM], __hip_atomic_load(&B.wq[k], __ATOMIC_RELAXED,
                                                                          __HIP_MEMORY_SCOPE_SYSTEM));
                            }
                            atomicOr(&B.ctl[FLOOD_OVF], OVF_TIME);
                            stop = 1;
                            break;
                        }
                    }
                    // an idle workgroup polls its slot every ~256 cycles at first; with backoff it then
                    // sleeps longer, leaving the CU's issue slots and the memory fabric to working waves
                    if (backoff == 0 || it < 32u) __builtin_amdgcn_s_sleep(4);
                    else if (backoff == 1 || it < 512u) __builtin_amdgcn_s_sleep(32);
                    else __builtin_amdgcn_s_sleep(127);
                }
                // leave the slot clean for the next call, with the same (system) scope as the pushes and
                // polls of queue slots: a write of another scope may reach memory in another order
                if (item != Q_EMPTY) __hip_atomic_store(&B.wq[k], Q_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            s_item = item;
            s_stop = stop;
            if (tprof) {
                const uint64_t now = __builtin_amdgcn_s_memrealtime();
                atomicAdd(&B.ctl[PROF_WAIT], uint32_t(now - tw0));
                if (stop) atomicAdd(&B.ctl[PROF_LIFE], uint32_t(now - t_born));
            }
        }
        __syncthreads();
        if (s_stop) break;
        const uint32_t ti = s_item;
        Tile &T = B.tiles[ti];
        if (tid == 0) {
            atomicExch(&T.qstate, 1u);   // in process: touches from now on ask for another pass
            atomicAdd(&B.ctl[Q_ITEMS], 1u);
            s_evald = ald(&T.evald);
            s_hint = ald(&T.hint);
        }
        if (tid < FT) {
            s_pass[tid] = ald(&T.pass[tid]);
            s_cnt[tid] = ald(&T.cnt[tid]);
            s_hot[tid] = ald(&T.hot[tid]);
        }
        __syncthreads();
        const FloodJob J = sa.nst ? job_load(&jobs[ald(&T.job)]) : jobs[ald(&T.job)];
        const int tx = ald(&T.tx), tz = ald(&T.tz);
        const uint32_t evald0 = s_evald;
        for (;;) {
            // ---- evaluate the touched sub-blocks until the reached set is final, then propagate
            for (;;) {
                if (tid < 64) {
                    const uint32_t need = prop_tile(B, ti, J, s_pass, s_evald, s_hint);
                    if (tid == 0) s_need = need;
                }
                __syncthreads();
                uint32_t need = s_need;
                if (!need) break;
                if (tid == 0) {
                    atomicAdd(&B.ctl[EVAL_ROUNDS], 1u);
                    atomicAdd(&B.ctl[EVAL_NEED], uint32_t(__popc(need)));
                }
                if (spec) {
                    // waves left idle by this round (the needed count is not a multiple of the waves) take
                    // sub-blocks next to the needed ones that are not evaluated yet, inside the window: the
                    // flood usually grows into them next, and evaluated cells are only a cache of the
                    // predicate (what is reached and counted still follows the touched cells alone)
                    const int slots = pairs ? NWAVE / 2 : NWAVE;   // sub-blocks one evaluation pass takes
                    int idle = (slots - __popc(need) % slots) % slots;
                    if (idle) {
                        const int nw = int(nwin(J));
                        uint32_t nb = 0;
                        for (uint32_t m = need; m; m &= m - 1) {
                            const int b = __ffs(m) - 1, bx = b % NSB, bz = b / NSB;
                            if (bx > 0) nb |= 1u << (b - 1);
                            if (bx < NSB - 1) nb |= 1u << (b + 1);
                            if (bz > 0) nb |= 1u << (b - NSB);
                            if (bz < NSB - 1) nb |= 1u << (b + NSB);
                        }
                        nb &= ~(need | s_evald);
                        for (; idle && nb; nb &= nb - 1) {
                            const int b = __ffs(nb) - 1;
                            if (tx * FT + (b % NSB) * SB < nw && tz * FT + (b / NSB) * SB < nw) {
                                need |= 1u << b;
                                idle--;
                            }
                        }
                    }
                }
                const uint64_t te0 = tprof && tid == 0 ? __builtin_amdgcn_s_memrealtime() : 0;
                const int cls = J.kind == PRED_HOT ? 0 : (J.kind == PRED_PROXY ? 1 : 2);
                if (J.seed != loaded_seed || cls > loaded_cls) {   // the octaves the predicate reads, in LDS
                    const ClimSeed &G = cs[J.seed];
                    constexpr int OW = sizeof(ClimOct) / 16;
                    if (cls < 2) {
                        const int nslot = cls == 0 ? 4 : 14;
                        for (int k = tid; k < nslot * OW; k += NT) {
                            const int slot = PROXY_SLOTS[k / OW];   // ClimP holds them in this order
                            reinterpret_cast<uint4 *>(&C.oct[LIGHT ? k / OW : slot])[k % OW] =
                                reinterpret_cast<const uint4 *>(&G.oct[slot])[k % OW];
                        }
                        constexpr int GW = sizeof(GradTab) / 16;
                        if (tid < GW) reinterpret_cast<uint4 *>(&C.g)[tid] = reinterpret_cast<const uint4 *>(&G.g)[tid];
                        if (tid < 6) C.factor[tid] = G.factor[tid];
                    } else if constexpr (!LIGHT) {   // light batches hold no biome-predicate jobs
                        const uint4 *src = reinterpret_cast<const uint4 *>(&G);
                        uint4 *dst = reinterpret_cast<uint4 *>(&C);
                        for (int k = tid; k < int(sizeof(ClimSeed) / 16); k += NT) dst[k] = src[k];
                    }
                    __syncthreads();
                    loaded_seed = J.seed;
                    loaded_cls = cls;
                    if (tprof && tid == 0) atomicAdd(&B.ctl[PROF_LOAD], uint32_t(__builtin_amdgcn_s_memrealtime() - te0));
                }
                if constexpr (PAIRS) {
                    if (pairs) {
                        // pass q: the (2q + p)-th needed sub-block to pair p (both of its waves), every pair
                        // through the same number of passes
                        const int np = NWAVE / 2, pair = wave >> 1, side = wave & 1, nn = __popc(need);
                        uint32_t m = need;
                        for (int q = 0; q * np < nn; q++) {
                            int sb = -1;
                            for (int k = 0; k < np && m; k++, m &= m - 1)
                                if (k == pair) sb = __ffs(m) - 1;
                            eval_pair(sb, side, tx, tz, J, C, s_x[pair], s_pass, s_cnt, s_hot);
                        }
                    }
                }
                // the k-th needed sub-block goes to wave k % NWAVE
                if (!pairs) {
                    int k = 0;
                    for (uint32_t m = need; m; m &= m - 1, k++)
                        if (k % NWAVE == wave) eval_subblock<PRE, LIGHT>(__ffs(m) - 1, tx, tz, J, C, g, s_pass, s_cnt, s_hot);
                }
                __syncthreads();
                if (tid == 0) {
                    s_evald |= need;
                    if (tprof) atomicAdd(&B.ctl[PROF_EVAL], uint32_t(__builtin_amdgcn_s_memrealtime() - te0));
                }
                __syncthreads();
            }
            if (tid < FT && s_evald != evald0) {   // newly evaluated cells: store the tile's masks
                ast(&T.pass[tid], s_pass[tid]);
                ast(&T.cnt[tid], s_cnt[tid]);
                ast(&T.hot[tid], s_hot[tid]);
            }
            if (tid == 0) {
                if (s_evald != evald0) ast(&T.evald, s_evald);
                mem_done();
            }
            __syncthreads();
            if (tid == 0) {
                mem_done();
                uint32_t again = 0;
                if (atomicCAS(&T.qstate, 1u, 0u) == 2u) {   // touched while in process
                    atomicExch(&T.qstate, 1u);
                    again = 1;
                } else {
                    atomicSub(&B.ctl[Q_PENDING], 1u);        // this item is done (its pushes are counted)
                }
                s_again = again;
            }
            __syncthreads();
            if (!s_again) break;
        }
    }
}

// Per-tile reduction of the flood outputs into res[job]; leaves the tile and its hash slot clean.
__device__ void acc_tile(const FloodJob *jobs, FloodBufs B, uint32_t b, FloodRes *res, TileMask *masks, uint32_t *mcount,
                         uint32_t mcap) {
    Tile &T = B.tiles[b];
    const uint32_t job = ald(&T.job);
    const int tx = ald(&T.tx), tz = ald(&T.tz);
    const FloodJob J = jobs[job];
    const int lane = threadIdx.x, r = lane & 31;
    const bool c8 = (J.conn8 & 1) != 0;
    const uint32_t R = lane < FT ? ald(&T.reach[r]) : 0u, I = lane < FT ? ald(&T.in[r]) : 0u;
    const uint32_t up = __shfl(R, (r + 31) & 31, 32), dn = __shfl(R, (r + 1) & 31, 32);
    const uint32_t upv = r > 0 ? up : 0u, dnv = r < FT - 1 ? dn : 0u;
    uint32_t touched = R | I | (R << 1) | (R >> 1) | upv | dnv;
    if (c8) touched |= (upv << 1) | (upv >> 1) | (dnv << 1) | (dnv >> 1);
    const int nw = int(nwin(J));
    const int j = tz * FT + r;
    uint32_t win = (lane < FT && j < nw) ? col_mask(J, tx) : 0u;
    touched &= win;
    // world mask of this row
    const int z = J.qz + (j - J.half) * J.step;
    uint32_t world = 0;
    if (z >= -WORLD_Q && z <= WORLD_Q) {
        for (int c = 0; c < FT; c++) {
            const int x = J.qx + (tx * FT + c - J.half) * J.step;
            if (x >= -WORLD_Q && x <= WORLD_Q) world |= 1u << c;
        }
    }
    unsigned long long sampled = __popc(touched & world);
    unsigned long long hotc = __popc(touched & world & (lane < FT ? ald(&T.hot[r]) : 0u));
    const bool border = (touched & ~world) != 0;
    const uint32_t RC = R & (lane < FT ? ald(&T.cnt[r]) : 0u);
    unsigned long long cells = __popc(RC);
    int minx = INT_MAX, maxx = INT_MIN, minz = INT_MAX, maxz = INT_MIN;
    if (RC) {
        const int c0 = __ffs(RC) - 1, c1 = 31 - __clz(RC);
        minx = J.qx + (tx * FT + c0 - J.half) * J.step;
        maxx = J.qx + (tx * FT + c1 - J.half) * J.step;
        minz = maxz = z;
    }
    // a reached cell on the window boundary has a neighbour outside the window
    bool edge = false;
    if (R) {
        if (j == 0 || j == nw - 1) edge = true;
        if (tx == 0 && (R & 1u)) edge = true;
        const int last = nw - 1 - tx * FT;
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
        FloodRes &o = res[job];
        atomicAdd(&o.tiles, 1ull);
        atomicAdd(&o.evald, 64ull * __popc(ald(&T.evald)));
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
        mask_export(B, masks, mcap, k, lane, RC, job, tx, tz);
    }
    // leave the tile and its hash slot clean for the next call (everything above has read them)
    __builtin_amdgcn_wave_barrier();
    if (lane < FT) {
        ast(&T.reach[lane], 0u);
        aset(&T.in[lane], 0u);
        ast(&T.pass[lane], 0u);
        ast(&T.cnt[lane], 0u);
        ast(&T.hot[lane], 0u);
    }
    if (lane == 0) {
        const uint32_t slot1 = ald(&T.slot1);
        if (slot1) aset(&B.hkeys[slot1 - 1], HEMPTY);
        ast(&T.slot1, 0u);
        aset(&T.qstate, 0u);
        ast(&T.evald, 0u);
        aset(&T.hint, 0u);
    }
}

__global__ __launch_bounds__(64) void k_flood_acc(const FloodJob *jobs, FloodBufs B, FloodRes *res, TileMask *masks,
                                                  uint32_t mcap) {
    const uint32_t ntiles = min(ald(B.ntiles), B.tile_cap);
    for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) acc_tile(jobs, B, b, res, masks, &B.ctl[FLOOD_MASKS], mcap);
}

// Results, control words and (masks non-null) masks into pinned memory (t0 of nt threads), the control
// block of the batch two after this one zeroed. Reads only agent-scope state that the accumulation has
// completed.
__device__ void publish(const FloodRes *res, uint32_t n, const FloodBufs &B, const TileMask *masks, FloodRes *h_res,
                        uint32_t *h_ctl, TileMask *h_masks, uint32_t hmcap, uint32_t t0, uint32_t nt) {
    if (t0 < CTL_WORDS) {
        h_ctl[t0] = ald(&B.ctl[t0]);
        aset(&B.ctl_next[t0], 0u);
    }
    constexpr uint32_t RW = sizeof(FloodRes) / 4;
    const uint32_t *rs = reinterpret_cast<const uint32_t *>(res);
    uint32_t *rd = reinterpret_cast<uint32_t *>(h_res);
    for (uint32_t q = t0; q < n * RW; q += nt) rd[q] = ald(&rs[q]);
    if (!masks) return;
    const uint32_t nm = min(ald(&B.ctl[FLOOD_MASKS]), hmcap);
    constexpr uint32_t W = sizeof(TileMask) / 4;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(masks);
    uint32_t *dst = reinterpret_cast<uint32_t *>(h_masks);
    for (uint32_t q = t0; q < nm * W; q += nt) dst[q] = ald(&src[q]);
}

__global__ __launch_bounds__(256) void k_flood_publish(const FloodRes *res, uint32_t n, FloodBufs B, const TileMask *masks,
                                                       FloodRes *h_res, uint32_t *h_ctl, TileMask *h_masks, uint32_t hmcap) {
    publish(res, n, B, masks, h_res, h_ctl, h_masks, hmcap, blockIdx.x * blockDim.x + threadIdx.x, blockDim.x * gridDim.x);
}

// k_flood_acc, then the results and control words published by the last workgroup to finish (one launch
// fewer per batch); the batch's masks go to pinned memory as they are exported (B.h_masks): each workgroup's
// accumulation (agent-scope atomics and stores) has completed (mem_done) before it counts itself out with
// an agent-scope atomic, so the workgroup that counts last reads every final result. No workgroup waits
// for another.
__global__ __launch_bounds__(64) void k_flood_acc_publish(const FloodJob *jobs, FloodBufs B, FloodRes *res, TileMask *masks,
                                                          uint32_t mcap, uint32_t n, FloodRes *h_res, uint32_t *h_ctl,
                                                          TileMask *h_masks, uint32_t hmcap) {
    const uint32_t ntiles = min(ald(B.ntiles), B.tile_cap);
    for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) acc_tile(jobs, B, b, res, masks, &B.ctl[FLOOD_MASKS], mcap);
    mem_done();
    uint32_t last = 0;
    if (threadIdx.x == 0) last = atomicAdd(&B.ctl[ACC_DONE], 1u) == gridDim.x - 1 ? 1u : 0u;
    if (!__shfl(last, 0)) return;
    publish(res, n, B, nullptr, h_res, h_ctl, nullptr, 0, threadIdx.x, blockDim.x);
}

// ------------------------------------------------------------------ launchers

void launch_flood_stage(hipStream_t st, const uint8_t *h_up, uint32_t n, uint32_t n_small, uint32_t n_big, uint32_t nh,
                        uint8_t *d_up, FloodRes *res, FloodBufs B) {
    const uint32_t m = std::max(n, nh);
    if (m) hipLaunchKernelGGL(k_flood_stage, dim3((m + 255) / 256), dim3(256), 0, st, h_up, n, n_small, n_big, nh, d_up, res, B);
}
void launch_flood_run(hipStream_t st, const FloodJob *jobs, const ClimSeed *cs, const SurfGrid &g, FloodBufs B,
                      uint32_t blocks, int pre, int light, hipEvent_t ev0, hipEvent_t ev1, int tprof,
                      const StageArgs *stage) {
    // workgroup size (DSX_FLOOD_NT = 64 to 1024 threads); `blocks` is given for 256-thread
    // workgroups and scaled so that the launch holds the same number of waves
    static const int nt_light = [] {
        const char *v = std::getenv("DSX_FLOOD_NT");
        const int n = v && *v ? std::atoi(v) : 256;
        return n == 64 || n == 128 || n == 512 || n == 1024 ? n : 256;
    }();
    // the full instance (batches with exact floods: 52 octave samples per hot cell, a latency-bound
    // chain of tiles) takes 512-thread workgroups, so that twice as many of a tile's sub-blocks are
    // evaluated at once (DSX_FLOOD_NT_FULL overrides; 1024 needs register spills and measured slower)
    static const int nt_full = [] {
        const char *v = std::getenv("DSX_FLOOD_NT_FULL");
        const int n = v && *v ? std::atoi(v) : 512;
        return n == 64 || n == 128 || n == 512 || n == 1024 ? n : 256;
    }();
    static const int occ = [] {   // DSX_FLOOD_OCC: 4 (both instances), 5 or 6 (the light one)
        const char *v = std::getenv("DSX_FLOOD_OCC");
        return v && *v ? std::atoi(v) : 1;
    }();
    const bool occ4 = occ >= 4;
    static const int backoff = [] {
        const char *v = std::getenv("DSX_FLOOD_BACKOFF");
        return v && *v ? std::atoi(v) : 0;
    }();
    static const int spec = [] {
        const char *v = std::getenv("DSX_FLOOD_SPEC");   // 0 turns the speculative sub-blocks off
        const char *p = std::getenv("DSX_FLOOD_PRIO");   // 1: flood waves at raised issue priority
        const char *w = std::getenv("DSX_FLOOD_PAIR");   // 0: one wave per sub-block in the light instance
        return (v && *v ? (std::atoi(v) != 0 ? 1 : 0) : 1) | (p && std::atoi(p) ? 2 : 0) | (w && *w == '0' ? 0 : 4);
    }();
    static const bool light_on = [] {
        const char *v = std::getenv("DSX_FLOOD_LIGHT");   // 0: always the full instance
        return !(v && *v == '0');
    }();
    light = light && light_on;
    const int nt = light ? nt_light : nt_full;
    decltype(&k_flood_run<256, 1, 0, 0>) kern;
#define DSX_FR(N, O) (pre ? (light ? k_flood_run<N, O, 1, 1> : k_flood_run<N, O, 1, 0>) \
                          : (light ? k_flood_run<N, O, 0, 1> : k_flood_run<N, O, 0, 0>))
    if (nt == 64) kern = DSX_FR(64, 1);
    else if (nt == 128) kern = DSX_FR(128, 1);
    else if (nt == 512) kern = DSX_FR(512, 1);
    else if (nt == 1024) kern = DSX_FR(1024, 1);
    else if (light && occ >= 5) kern = occ >= 6 ? (pre ? k_flood_run<256, 6, 1, 1> : k_flood_run<256, 6, 0, 1>)
                                                 : (pre ? k_flood_run<256, 5, 1, 1> : k_flood_run<256, 5, 0, 1>);
    else kern = occ4 ? DSX_FR(256, 4) : DSX_FR(256, 1);
#undef DSX_FR
    const dim3 grid(nt <= 256 ? (blocks ? blocks : 1) * (256 / nt) : std::max(1u, blocks * 256 / nt));
    StageArgs sa{};
    if (stage) {   // one staging workgroup per nt items, at most the whole grid
        sa = *stage;
        const uint32_t m = std::max(sa.n, std::max(sa.n_big, sa.nh));
        sa.nst = std::max(1u, std::min(grid.x, (m + uint32_t(nt) - 1) / uint32_t(nt)));
    }
    if (ev0) hipExtLaunchKernelGGL(kern, grid, dim3(nt), 0, st, ev0, ev1, 0u, jobs, cs, g, B, backoff, spec, tprof, sa);
    else hipLaunchKernelGGL(kern, grid, dim3(nt), 0, st, jobs, cs, g, B, backoff, spec, tprof, sa);
}
void launch_flood_acc(hipStream_t st, const FloodJob *jobs, FloodBufs B, FloodRes *res, TileMask *masks, uint32_t mcap,
                      uint32_t blocks) {
    hipLaunchKernelGGL(k_flood_acc, dim3(blocks ? blocks : 1), dim3(64), 0, st, jobs, B, res, masks, mcap);
}
void launch_flood_publish(hipStream_t st, const FloodRes *res, uint32_t n, FloodBufs B, const TileMask *masks,
                          FloodRes *h_res, uint32_t *h_ctl, TileMask *h_masks, uint32_t hmcap) {
    hipLaunchKernelGGL(k_flood_publish, dim3(64), dim3(256), 0, st, res, n, B, masks, h_res, h_ctl, h_masks, hmcap);
}
void launch_flood_acc_publish(hipStream_t st, const FloodJob *jobs, FloodBufs B, FloodRes *res, TileMask *masks,
                              uint32_t mcap, uint32_t blocks, uint32_t n, FloodRes *h_res, uint32_t *h_ctl,
                              TileMask *h_masks, uint32_t hmcap) {
    hipLaunchKernelGGL(k_flood_acc_publish, dim3(blocks ? blocks : 1), dim3(64), 0, st, jobs, B, res, masks, mcap, n, h_res,
                       h_ctl, h_masks, hmcap);
}

}  // namespace dsx
