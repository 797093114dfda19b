// Internal declarations shared by the HIP translation units of libdsx.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "dsx_rng.h"

namespace dsx {

constexpr int WORLD_Q = 7499996;   // world border in quarts (dense_flood.hpp:33, dsearch_desert.cpp:44)

// Temperature octave-0 Perlin of one noise side (the reference's tg::PerlinParams, gate.cuh:49).
struct SideParams {
    uint8_t perm[256];
    double ox, oz, fy, vy;
    int iy, pad_;
};

// Gate blur kernel as the kernels consume it: 2R+1 float weights, the float factor applied to the
// maximum, and the octave weight. GPU gate (gate.cuh:216, 248): unnormalised expf weights, factor
// 1/(ksum*ksum), weight float(1.25 * 32/63 * 1.5). CPU-default pre-gate (lowfield.hpp:92-96):
// normalised float(exp / sum) weights, factor 1, weight factor * octave weight in double.
// trunc: a bound, per unit of |grid value|, on how far the normalised blur moves when its outermost taps
// (d = +-R) are left out of both passes: norm * (ksum^2 - ksum_t^2), ksum_t the sum without them, rounded up
// (0: no truncated fast pre-gate for this kernel).
struct GateKernel {
    float k[15];
    float norm;
    float trunc;
    // The integer blur passes of the 16-bit fast amplitudes: the taps w they use (R = 4: taps 1 .. 7, the
    // truncated blur; R = 7: all 15) times SK (2^14 and 2^13), rounded, as 16-bit pairs: with NP = 4 or 8,
    // kq2[m] = (w[2m - 1], w[2m]) and kq2[NP + m] = (w[2m], w[2m + 1]), w[-1] = w[W] = 0. wq bounds the
    // normalised change that the integer passes (these weights, the horizontal sums rounded to 16 bits)
    // make per unit of |grid value|.
    uint32_t kq2[16];
    float wq;
    double wd;   // octave weight the corner tables are scaled by
};

// The gate as one chain of kernels that needs no host round trip (no counts read back between stages):
// counters cleared -> phase filter -> the count the later stages take (all phase survivors, or none when they
// exceed `grid`) -> Perlin init -> pre-gate -> full gate -> survivors rank-sorted by seed index into `out`.
// The grids of the later stages are sized for `grid` phase survivors; their kernels read the live counts.
// With fast set, the pre-gate first decides with the fast amplitude (FMA blurs, provably the same decisions)
// and runs the reference's arithmetic only where a decision is within its error bound (gate_launches in
// gate.hip); every amplitude output is the reference's.
constexpr int GATE_NCNT = 8;   // counters: phase, pre, full, taken, pre undecided, full undecided
struct GateChain {
    uint32_t *cnt;        // GATE_NCNT counters
    uint32_t *k1, *i2, *k3;
    float *a3;
    SideParams *sides;    // 2 * grid
    uint32_t grid;        // phase survivors the chain holds
    bool fast = false;
    uint32_t *u2 = nullptr;   // (GateStages::full_fast) the full gate's undecided pairs, grid entries
    float *band = nullptr;    // (fast chains) band maxima of the undecided pairs, grid * 16 entries
};
struct GateOut {          // device or pinned host memory the sort kernel writes
    uint64_t *seeds;      // may be null
    float *amps;          // may be null
    uint64_t *kidx;       // may be null
    uint64_t cap;         // entries written at most
    uint64_t *counts;     // {phase, pre, full (0, 0 if the phase survivors overflowed), min(full, cap)}
};
struct GateStages {       // thresholds and kernels of one gate configuration
    double maxph, thr_pre, amp;
    int mode_pre;         // 0 anti-aliased pre-gate, 2 legacy even-parity pre-gate
    int full_fast;        // (fast chains) the full gate decided from its fast amplitude: survivors only, no
                          // amplitudes for the seeds it decides
    GateKernel kpre, kfull;
};
// Enqueues the chain on st for seed indices [start, start + count) (count <= 2^31).
hipError_t enqueue_gate_chain(hipStream_t st, const GateChain &C, const GateStages &S, uint64_t start, uint32_t count,
                              const GateOut &O);
// The same chain as a hipGraph of kernel nodes (no memset nodes), built once per context and configuration and
// replayed per chunk with only its start and count changed (count <= nmax, the size it was built for).
struct GateGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t n_phase = nullptr, n_init = nullptr, n_sort = nullptr;
    GateChain C{};
    GateStages S{};
    GateOut O{};
    uint32_t nmax = 0;
};
hipError_t gate_graph_build(GateGraph &G, const GateChain &C, const GateStages &S, uint32_t nmax, const GateOut &O);
hipError_t gate_graph_launch(GateGraph &G, hipStream_t st, uint64_t start, uint32_t count);
void gate_graph_free(GateGraph &G);

void launch_phase(hipStream_t st, uint64_t base, uint32_t n, double maxph, uint32_t *out, uint32_t *count, uint32_t cap);
// n_dev (optional): a device word holding the item count, read by the kernels (min(n, *n_dev)); the grid
// still covers n items, so a caller that does not know the count on the host launches with an upper bound.
void launch_init(hipStream_t st, uint64_t base, const uint32_t *ks, const uint64_t *seeds, uint32_t n, SideParams *out,
                 const uint32_t *n_dev = nullptr);
// Amplitude kernel k_amp<mode, dec> (gate.hip): dec = GATE_EXACT, or GATE_SPLIT (modes 0 and 2), whose
// undecided list is und_idx / und_count, or GATE_SPLIT_T (mode 0): GATE_SPLIT with the blur's outermost taps
// left out of the fast amplitude, its decision margin widened by GateKernel::trunc, or GATE_SPLIT_Q (modes 0
// and 1): the grid summed in 16-bit fixed point and blurred in 16-bit integer dot products (mode 0 also
// truncated as GATE_SPLIT_T), the margin widened by the bounds of both.
enum { GATE_EXACT = 0, GATE_SPLIT = 1, GATE_SPLIT_T = 2, GATE_SPLIT_Q = 3 };
// In a chain's launch list (GateLaunch::dec) only: the reference's arithmetic over a short list of undecided
// pairs, one wave per (pair, side, band of 64 output rows) and then a kernel that takes each side's maximum
// over its bands and decides (the same values and the same decisions as GATE_EXACT, whose one block per
// pair would keep a CU busy for the whole 256^2 or 512^2 grid); needs GateChain::band.
enum { GATE_EXACT_BANDS = 4 };
void launch_amp(hipStream_t st, int mode, int dec, const SideParams *sides, const uint32_t *sel, uint32_t n,
                const GateKernel &K, double thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp, uint32_t *count,
                uint32_t cap, float *all_amp, const uint32_t *n_dev = nullptr, uint32_t *und_idx = nullptr,
                uint32_t *und_count = nullptr);
// One amplitude stage of a chain: k_amp<mode, dec> over sel (null: all phase survivors), its count in *ndev,
// with the pre-gate (pre) or full-gate kernel and threshold.
struct GateLaunch {
    int mode, dec;
    bool pre;
    const uint32_t *sel, *kmap;
    uint32_t *out;
    float *amp;
    uint32_t *count;
    const uint32_t *ndev;
    uint32_t *und, *und_count;
};

}  // namespace dsx
