// Floods in one wave with their state in LDS (dense_flood.hpp:52-86), evaluating exactly the cells
// the reference touches.
//
// The wave keeps a local sub-window of at most 255 x 255 lattice points around the flood's start: the
// whole window when it is that small (the step-256 hot-zone pre-floods of dsearch_desert.cpp:337),
// else the 255 x 255 points centred on the start (step-128 hot zones :352 and step-32 proxy floods
// :387, whose components are almost always far smaller than their windows). The tiled engine would
// evaluate whole 32x32 tiles around such components; here:
//   * touched and counted bitmaps (1 bit per local cell each) and a ring of pushed cells live in LDS;
//   * a step pops up to 64 pushed cells (one per lane), claims their untouched in-window neighbours
//     with LDS atomicOr on the touched bitmap and lists the claimed cells; the listed cells are then
//     evaluated 64 at a time (all lanes busy on the FP64 predicate), and passable ones are pushed;
//   * the outputs (cells, sampled, bbox, edge, border) are order independent, so the visiting order
//     (wave-parallel, ring order) gives exactly the reference's depth-first results.
// A flood that reaches the edge of its local sub-window inside the real window, or whose ring would
// overflow, is handed to the tiled engine (appended to its job list, which runs next on the stream).
#include "floods.h"
#include "flood_dev.h"
#include <climits>
#include <type_traits>

namespace dsx {

// Local sub-window edge N (255, or 64 for threshold floods): bitmap words and pushed cells in flight.
template <int N> struct LdsShape {
    static constexpr int WORDS = (N * N + 31) / 32;   // 2033 words per bitmap at N = 255
    static constexpr int RING = N >= 255 ? 2048 : 512;
};

template <class CS, int PRE>
__device__ __forceinline__ int lds_pred(int kind, const CS &C, const SurfGrid &g, int x, int z, float cut, bool *hot) {
    if constexpr (std::is_same<CS, ClimT>::value) {
        return PRE ? hot_state_pre(C, x, z) : hot_state(C, x, z);
    } else if constexpr (std::is_same<CS, ClimP>::value) {
        return PRE ? proxy_state_pre(C, x, z, cut, hot) : proxy_state(C, x, z, cut, hot);
    } else {
        return PRE ? predicate_pre(kind, C, g, x, z, cut, hot) : predicate(kind, C, g, x, z, cut, hot);
    }
}

// 32 bits of row lj of a local bitmap (row length L bits) starting at local column c0 (c0 may be
// negative or run past L: those columns read as 0).
__device__ __forceinline__ uint32_t bitmap_row32(const uint32_t *bm, int words, int L, int lj, int c0) {
    const int lo = c0 < 0 ? 0 : c0, hi = c0 + 32 > L ? L : c0 + 32;   // columns [lo, hi) exist
    if (hi <= lo) return 0u;
    const int b = lj * L + lo;
    const int w0 = b >> 5, sh = b & 31;
    const uint64_t two = uint64_t(bm[w0]) | (uint64_t(w0 + 1 < words ? bm[w0 + 1] : 0u) << 32);
    uint32_t v = uint32_t(two >> sh);
    const int n = hi - lo;
    if (n < 32) v &= (1u << n) - 1u;
    return v << (lo - c0);
}

template <class CS, int PRE, int N>
__global__ __launch_bounds__(64) void k_flood_lds(const FloodJob *jobs, const uint32_t *list, uint32_t n, const ClimSeed *cs,
                                                  SurfGrid g, FloodRes *res, TileMask *masks, uint32_t mcap, FloodBufs B) {
    uint32_t *mcount = B.ctl + FLOOD_MASKS;
    __shared__ CS C;
    constexpr int LDS_N = N, LDS_WORDS = LdsShape<N>::WORDS, LDS_RING = LdsShape<N>::RING;
    __shared__ uint32_t touched[LDS_WORDS], counted[LDS_WORDS];
    __shared__ uint32_t ring[LDS_RING];
    __shared__ uint32_t evl[512];
    __shared__ uint32_t qhead, qtail, nev, ovf, ncnt;
    const uint32_t b = blockIdx.x;
    if (b >= n) return;
    const uint32_t job = list[b];
    const FloodJob J = jobs[job];
    const int lane = threadIdx.x;
    const int nw = 2 * J.half + 1;
    // local sub-window: columns/rows [o, o + L) of the window
    const int L = nw < LDS_N ? nw : LDS_N, o = J.half - (L - 1) / 2;
    const int ncell = L * L;
    {   // climate (only the octaves the predicate reads unless CS is ClimSeed) and cleared bitmaps
        copy_climate(C, cs[J.seed], lane, 64);
        const int words = (ncell + 31) / 32;
        for (int k = lane; k < words; k += 64) {
            touched[k] = 0;
            counted[k] = 0;
        }
        if (lane == 0) {
            qhead = 0;
            qtail = 0;
            nev = 0;
            ovf = 0;
            ncnt = 0;
        }
    }
    __syncthreads();
    unsigned long long cells = 0, sampled = 0, hotc = 0;
    int minx = J.qx, maxx = J.qx, minz = J.qz, maxz = J.qz;
    const uint32_t stop_at = uint32_t(J.conn8) >> 8;
    bool capped = false;
    bool edge = false, border = false;
    // the centre is touched first (dense_flood.hpp:80)
    if (lane == 0) {
        const int c = (J.half - o) * L + (J.half - o);
        touched[c >> 5] |= 1u << (c & 31);
        evl[0] = uint32_t(c);
        nev = 1;
    }
    __syncthreads();
    for (;;) {
        // ---- evaluate the claimed cells, push the passable ones
        const uint32_t ne = nev;
        for (uint32_t e = lane; e < ne; e += 64) {
            const int c = int(evl[e]), i = c % L + o, j = c / L + o;
            const int x = J.qx + (i - J.half) * J.step, z = J.qz + (j - J.half) * J.step;
            if (x < -WORLD_Q || x > WORLD_Q || z < -WORLD_Q || z > WORLD_Q) {
                border = true;
                continue;
            }
            sampled++;
            bool hot = false;
            const int v = lds_pred<CS, PRE>(J.kind, C, g, x, z, J.cut, &hot);
            hotc += hot;
            if (v == 0) continue;
            if (v == 2) {
                atomicOr(&counted[c >> 5], 1u << (c & 31));
                atomicAdd(&ncnt, 1u);
                cells++;
                minx = min(minx, x);
                maxx = max(maxx, x);
                minz = min(minz, z);
                maxz = max(maxz, z);
            }
            const uint32_t t = atomicAdd(&qtail, 1u);
            if (t - qhead >= LDS_RING) ovf = 1;
            else ring[t % LDS_RING] = uint32_t(c);
        }
        __syncthreads();
        if (lane == 0) nev = 0;
        const uint32_t h = qhead, tl = qtail;
        // a threshold flood (job stop_cells) is done once it has counted that many cells
        const bool stop = stop_at != 0u && ncnt >= stop_at;
        __syncthreads();
        if (ovf || h == tl || stop) {
            capped = stop && !ovf && h != tl;
            break;
        }
        // ---- expand up to 64 pushed cells: claim their untouched neighbours
        const uint32_t np = min(tl - h, 64u);
        if (uint32_t(lane) < np) {
            const int c = int(ring[(h + lane) % LDS_RING]), li = c % L, lj = c / L;
            const int nn = (J.conn8 & 1) ? 8 : 4;
            for (int d = 0; d < nn; d++) {
                const int di = d == 0 ? -1 : d == 1 ? 1 : d == 2 ? 0 : d == 3 ? 0 : d == 4 ? -1 : d == 5 ? 1 : d == 6 ? -1 : 1;
                const int dj = d == 0 ? 0 : d == 1 ? 0 : d == 2 ? -1 : d == 3 ? 1 : d == 4 ? -1 : d == 5 ? -1 : d == 6 ? 1 : 1;
                const int ni = li + di + o, nj = lj + dj + o;   // window coordinates
                if (ni < 0 || nj < 0 || ni >= nw || nj >= nw) {
                    edge = true;
                    continue;
                }
                const int qi = ni - o, qj = nj - o;
                if (qi < 0 || qj < 0 || qi >= L || qj >= L) {   // beyond the local sub-window
                    ovf = 1;
                    continue;
                }
                const int q = qj * L + qi;
                const uint32_t bit = 1u << (q & 31);
                if (atomicOr(&touched[q >> 5], bit) & bit) continue;
                evl[atomicAdd(&nev, 1u)] = uint32_t(q);
            }
        }
        __syncthreads();
        if (lane == 0) qhead = h + np;
        __syncthreads();
        if (ovf) break;
    }
    // ---- reduce
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
        cells += __shfl_xor(cells, o2);
        sampled += __shfl_xor(sampled, o2);
        hotc += __shfl_xor(hotc, o2);
        minx = min(minx, __shfl_xor(minx, o2));
        maxx = max(maxx, __shfl_xor(maxx, o2));
        minz = min(minz, __shfl_xor(minz, o2));
        maxz = max(maxz, __shfl_xor(maxz, o2));
    }
    const bool any_edge = __any(edge), any_border = __any(border);
    if (ovf) {   // not done here: hand it to the tiled engine
        if (lane == 0) {
            FloodRes r{};
            r.minx = r.maxx = J.qx;
            r.minz = r.maxz = J.qz;
            res_store(&res[job], r);
            atomicAdd(&B.ctl[HANDED], 1u);
            flood_start_tile(J, job, B);
        }
        return;
    }
    if (lane == 0) {
        FloodRes r;
        r.cells = cells;
        r.sampled = sampled;
        r.minx = minx;
        r.maxx = maxx;
        r.minz = minz;
        r.maxz = maxz;
        r.edge = any_edge;
        r.border = any_border;
        r.hot = hotc;
        r.tiles = 0;
        r.evald = sampled;
        res_store(&res[job], r);
    }
    if (cells == 0 || capped) return;   // a stopped flood's cells are not exported
    // ---- counted cells as 32x32 tile masks (window tile coordinates) over the counted bounding box
    const int i0 = (minx - J.qx) / J.step + J.half, i1 = (maxx - J.qx) / J.step + J.half;
    const int j0 = (minz - J.qz) / J.step + J.half, j1 = (maxz - J.qz) / J.step + J.half;
    for (int tz = j0 >> 5; tz <= (j1 >> 5); tz++)
        for (int tx = i0 >> 5; tx <= (i1 >> 5); tx++) {
            uint32_t row = 0;
            if (lane < 32) {
                const int j = tz * 32 + lane;                 // window row
                if (j < nw && j - o >= 0 && j - o < L) {
                    row = bitmap_row32(counted, LDS_WORDS, L, j - o, tx * 32 - o);
                    const int k = nw - tx * 32;               // valid columns of this row
                    if (k < 32) row &= (1u << k) - 1u;
                }
            }
            if (!__any(row != 0)) continue;
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(mcount, 1u);
            k = __shfl(k, 0);
            mask_export(B, masks, mcap, k, lane, row, job, tx, tz);
        }
}

template <int PRE>
static void lds_launch(hipStream_t st, int cls, const FloodJob *jobs, const uint32_t *list, uint32_t n, const ClimSeed *cs,
                       const SurfGrid &g, FloodRes *res, TileMask *masks, uint32_t mcap, FloodBufs B) {
    if (cls == LDS_PROXY_SMALL)
        hipLaunchKernelGGL((k_flood_lds<ClimP, PRE, 64>), dim3(n), dim3(64), 0, st, jobs, list, n, cs, g, res, masks, mcap, B);
    else if (cls == LDS_HOT_SMALL)
        hipLaunchKernelGGL((k_flood_lds<ClimT, PRE, 64>), dim3(n), dim3(64), 0, st, jobs, list, n, cs, g, res, masks, mcap, B);
    else if (cls == LDS_HOT)
        hipLaunchKernelGGL((k_flood_lds<ClimT, PRE, 255>), dim3(n), dim3(64), 0, st, jobs, list, n, cs, g, res, masks, mcap, B);
    else if (cls == LDS_PROXY)
        hipLaunchKernelGGL((k_flood_lds<ClimP, PRE, 255>), dim3(n), dim3(64), 0, st, jobs, list, n, cs, g, res, masks, mcap,
                           B);
    else
        hipLaunchKernelGGL((k_flood_lds<ClimSeed, PRE, 255>), dim3(n), dim3(64), 0, st, jobs, list, n, cs, g, res, masks,
                           mcap, B);
}

void launch_flood_lds(hipStream_t st, int cls, const FloodJob *jobs, const uint32_t *list, uint32_t n, const ClimSeed *cs,
                      const SurfGrid &g, FloodRes *res, TileMask *masks, uint32_t mcap, FloodBufs B, int pre) {
    if (!n) return;
    if (pre) lds_launch<1>(st, cls, jobs, list, n, cs, g, res, masks, mcap, B);
    else lds_launch<0>(st, cls, jobs, list, n, cs, g, res, masks, mcap, B);
}

}  // namespace dsx
