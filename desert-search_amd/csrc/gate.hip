// Seed gate kernels for gfx950 (wave64): y-phase filter -> Perlin init -> amplitude gates.
//
// Semantics follow the reference's GPU gate, src/gpu/gate.cuh:
//   k_phase   gate.cuh:87-92   seed = splitmix64(base + k); keep k if phase(seed) <= (double)maxph
//   k_init    gate.cuh:93-97   temperature octave-0 Perlin (perm + offsets) of both noise sides
//   k_amp<M>  gate.cuh:205-346 per side: value grid -> periodic horizontal blur -> vertical blur ->
//                              maximum * 1/ksum^2; seed passes if maxA + maxB >= thr
// with the same float operations in the same order as gate.cuh, so amplitudes are bit-identical to a
// CPU evaluation of that arithmetic (the CPU oracle). Design differs from the reference throughout:
//   * survivors are compacted with a wave ballot + one atomic per wave (no per-thread atomics);
//   * the permutation shuffle runs one lane per noise side with the 256-entry table in a padded,
//     lane-private LDS slot, then a transposed, coalesced write-out;
//   * the amplitude kernels use one wave per noise side and NO block-wide barrier in the main loop:
//     each lane owns G/64 adjacent grid columns and walks all rows; the horizontal blur takes its
//     halo columns from the neighbouring lanes by cross-lane shuffle (periodic: lane 63 wraps to 0),
//     and the vertical blur keeps its 2R+1-row window in registers. Only the read-only corner tables
//     and hash streams live in LDS, laid out so every table read of a wave hits one 16-dword
//     sub-table (conflict-free).
#include "grid_blur.h"
#include <climits>
#include <cstdlib>
#include <vector>

namespace dsx {

// ------------------------------------------------------------------ y-phase filter
// Each 256-thread workgroup takes PHASE_PER_WG = 2048 consecutive indices (8 per thread, coalesced) and
// reserves its survivors' slots with one atomic: some 15 survivors per workgroup at max_phase 0.06. (One
// atomic per wave that has a survivor, ~97 k per 2^24 indices on one counter, kept the kernel at 1.1 ms:
// the atomics on one address serialise; the arithmetic takes ~0.1 ms.) Slots are in any order; the chain
// ranks its survivors by index later.
constexpr uint32_t PHASE_PER_THREAD = 8, PHASE_PER_WG = 256 * PHASE_PER_THREAD;
__global__ __launch_bounds__(256) void k_phase(uint64_t base, uint32_t n, double maxph, uint32_t *out, uint32_t *count,
                                               uint32_t cap) {
    __shared__ uint32_t s_wtot[4], s_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t k0 = blockIdx.x * PHASE_PER_WG + threadIdx.x;
    uint64_t bal[PHASE_PER_THREAD];
    uint32_t wtot = 0;
#pragma unroll
    for (uint32_t i = 0; i < PHASE_PER_THREAD; i++) {
        const uint32_t k = k0 + i * 256u;
        const bool pass = k < n && temperature_phase(mix64(base + k), nullptr) <= maxph;
        bal[i] = __ballot(pass);
        wtot += uint32_t(__popcll(bal[i]));
    }
    if (lane == 0) s_wtot[wave] = wtot;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = s_wtot[0] + s_wtot[1] + s_wtot[2] + s_wtot[3];
        s_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    if (wtot == 0) return;
    uint32_t slot = s_base;
    for (int w = 0; w < wave; w++) slot += s_wtot[w];
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (uint32_t i = 0; i < PHASE_PER_THREAD; i++) {
        if ((bal[i] >> lane) & 1ull) {
            const uint32_t idx = slot + uint32_t(__popcll(bal[i] & lt));
            if (idx < cap) out[idx] = k0 + i * 256u;
        }
        slot += uint32_t(__popcll(bal[i]));
    }
}

// ------------------------------------------------------------------ Perlin init, one lane per side
constexpr int PERM_STRIDE = 260;   // 65 dwords: lane-private slots start in different LDS banks

// n_dev (optional): the number of seeds is min(n, *n_dev), read on the device (the grid covers n).
__global__ __launch_bounds__(64) void k_init(uint64_t base, const uint32_t *ks, const uint64_t *seeds, uint32_t n,
                                             SideParams *out, const uint32_t *n_dev) {
    __shared__ uint32_t slot32[64 * PERM_STRIDE / 4];
    uint8_t *slots = reinterpret_cast<uint8_t *>(slot32);
    if (n_dev) n = min(n, *n_dev);
    if (blockIdx.x * 64u >= 2u * n) return;
    const int lane = threadIdx.x;
    const uint32_t g = blockIdx.x * 64u + lane;
    const bool active = g < 2u * n;
    uint8_t *my = slots + lane * PERM_STRIDE;
    double ox = 0, oy = 0, oz = 0;
    if (active) {
        const uint64_t seed = seeds ? seeds[g >> 1] : mix64(base + ks[g >> 1]);
        Xoro c = temperature_fork(seed);
        uint64_t fa = c.next(), fb = c.next();
        if (g & 1) { fa = c.next(); fb = c.next(); }
        Xoro r = Xoro::raw(fa ^ DSX_OCT_SALT[0][0], fb ^ DSX_OCT_SALT[0][1]);
        ox = r.uniform() * 256;
        oy = r.uniform() * 256;
        oz = r.uniform() * 256;
        uint32_t *my32 = reinterpret_cast<uint32_t *>(my);
        for (int w = 0; w < 64; w++) my32[w] = 0x03020100u + 0x04040404u * uint32_t(w);
        for (int i = 0; i < 256; i++) {
            const int j = i + int(r.bounded(uint32_t(256 - i)));
            const uint8_t t = my[i];
            my[i] = my[j];
            my[j] = t;
        }
    }
    __syncthreads();
    const uint32_t g0 = blockIdx.x * 64u;
    for (int s = 0; s < 64; s++) {
        if (g0 + s >= 2u * n) break;
        reinterpret_cast<uint32_t *>(out[g0 + s].perm)[lane] = slot32[s * (PERM_STRIDE / 4) + lane];
    }
    if (active) {
        SideParams &p = out[g];
        p.ox = ox;
        p.oz = oz;
        p.fy = oy - floor(oy);
        p.vy = fade(p.fy);
        p.iy = int(floor(oy));
    }
}

// ------------------------------------------------------------------ amplitude gates
// MODE 0: anti-aliased 256^2 pre-gate (R=4); MODE 1: full 512^2 amplitude (R=7); MODE 2: legacy
// 256^2 even-parity pre-gate (R=4). Grid arithmetic and blur: grid_blur.h.
template <int MODE> struct GateR { static constexpr int R = (MODE == 1) ? 7 : 4; };

// A bound on |grid value| of one side (the corner tables must be set up): a value sums one entry of each
// sub-table of its cell's set (MODE 0: all 18; MODE 1: the 8 of one of the 4 parities; MODE 2: all 8), so
// the sum of those sub-tables' largest |entry| bounds it, up to the float roundings of both sums (covered
// by the margin of gate_eps).
template <int MODE>
__device__ float side_bound(const WaveLds &L) {
    constexpr int NSUB = MODE == 0 ? 18 : MODE == 1 ? 32 : 8;
    const int lane = threadIdx.x & 63;
    float m = 0.f;
    if (lane < NSUB)
        for (int h = 0; h < 16; h++) m = fmaxf(m, fabsf(L.tab[lane * 16 + h]));
    // sum over groups of 8 lanes (one parity's corner terms in MODE 1), then over (MODE 0, 2) or the
    // maximum of (MODE 1) the groups; lanes >= NSUB hold 0
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) m += __shfl_xor(m, o);
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) m = MODE == 1 ? fmaxf(m, __shfl_xor(m, o)) : m + __shfl_xor(m, o);
    return m;
}

// FAST: fused multiply-add blurs; TRUNC (with FAST): the blur's outermost taps left out (radius R - 1, taps
// 1 .. 2R - 1 of K.k), which moves the normalised result by at most K.trunc per unit of |grid value|.
// Q16 (with FAST; MODE 0 truncated to radius 3, or MODE 1 at radius 7): the grid in 16-bit fixed point
// (grid_setup_q16 into tq: MODE 0 packs it into L.tab, MODE 1 needs 768 words of its own) and both blur
// passes in integer dot products; b = the side's grid bound, which it stores in *bound.
template <int MODE, bool FAST, bool TRUNC = false, bool Q16 = false>
__device__ float side_max(const SideParams &sp, WaveLds &L, const GateKernel &K, float *bound = nullptr,
                          uint32_t *tq = nullptr) {
    static_assert(FAST || !TRUNC, "only the fast amplitude may leave taps out");
    static_assert(!Q16 || (FAST && ((MODE == 0 && TRUNC) || (MODE == 1 && !TRUNC))),
                  "the 16-bit forms: the truncated 256^2 grid, the full 512^2 grid");
    grid_setup<MODE>(sp, L, K.wd);
    float inv_s = 1.f;
    uint32_t qbias[4] = {0, 0, 0, 0};
    if constexpr (Q16) {
        const float b = side_bound<MODE>(L);
        *bound = b;
        if (MODE == 0) tq = reinterpret_cast<uint32_t *>(L.tab);
        inv_s = 1.f / grid_setup_q16<MODE>(L, b, tq, qbias);
    }
    float best = -1e9f;
    int ibest = INT_MIN;
    constexpr int NC = Shape<MODE>::G / 64, R = GateR<MODE>::R - (TRUNC ? 1 : 0);
    constexpr bool INTB = Q16;   // both blur passes in integers (grid_blur_rows, QH)
    // only the maximum is kept, so sums may start from their first term (ZSIGN off)
    grid_blur_rows<MODE, R, false, FAST, Q16>(sp, L, K.k + (TRUNC ? 1 : 0), 0, Shape<MODE>::G - 1,
                                              [&](int, const auto &o) {
#pragma unroll
        for (int q = 0; q < NC; q++) {
            if constexpr (INTB) ibest = max(ibest, o[q]);
            else best = fmaxf(best, o[q]);
        }
    }, K.kq2, tq, qbias);
    if constexpr (INTB) best = float(ibest);   // one rounding (inside gate_eps)
    // wave maximum; x -> fl(x * norm) is monotone, so max(s) * norm == max(s * norm)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o));
    // fixed point: the maximum in units of 2^16 / (SK^2 S) after the integer blur passes (SK = 2^14 at
    // radius 3, 2^13 at radius 7; two more roundings, a few ulps of |amp|: inside gate_eps)
    return Q16 ? best * (K.norm * (inv_s * (R == 3 ? 0x1p-12f : 0x1p-10f))) : best * K.norm;
}

// Decision margin of the fast amplitude (FMA blurs) for sides bounded by bA, bB. Each blur pass of W <= 15
// taps, fused or not, is within gamma_15 * sum_d k_d |input| of its exact value (gamma_15 = 15u / (1 - 15u)
// <= 9e-7, u = 2^-24), so the fused and the reference result of the two passes differ by at most
// 4 gamma_15 * ksum^2 * b before the factor norm = 1 / ksum^2, i.e. 3.6e-6 * b per side; the maximum over
// cells moves by no more than the largest difference, and the roundings of `* norm` and of the sum of the
// two sides add at most 4u |amp| < 1e-6 for |amp| < 4. 2e-5 * (bA + bB) + 1e-5 covers this 5 times over.
__device__ __forceinline__ double gate_eps(float bA, float bB) { return 2e-5 * (double(bA) + double(bB)) + 1e-5; }

// One 128-thread block per seed: wave 0 = side A, wave 1 = side B. sel (optional) maps block -> side
// pair index. n_dev (optional): the number of seeds is min(n, *n_dev), read on the device (the grid covers n).
// DEC selects what is decided and written:
//   GATE_EXACT  the reference's arithmetic: seeds with amplitude >= thr are appended as
//               (kmap ? kmap[pair] : pair, amplitude); all_amp (optional) gets every amplitude, by block
//   GATE_SPLIT  the fast amplitude (FMA blurs): pairs of seeds certainly at or above thr (amp - eps >= thr)
//               to out_idx, those within eps of thr to und_idx (to be decided by GATE_EXACT); all_amp
//               (optional) gets every fast amplitude and out_amp (optional, with all_amp) its margin eps
//   GATE_SPLIT_T the same with the blur's outermost taps left out (radius R - 1 in both passes: a quarter
//               fewer blur operations at R = 4), eps widened by K.trunc * (bA + bB)
//   GATE_SPLIT_Q the grid summed in 16-bit fixed point, two cells per add, and both blur passes in 16-bit
//               integer dot products: MODE 0 with GATE_SPLIT_T's truncation, eps widened by a further
//               (2.9e-4 + K.wq) * (bA + bB); MODE 1 at the full radius, eps widened by (1.3e-4 + K.wq) *
//               (bA + bB). Certain passes go out as GATE_EXACT's do (kmap ? kmap[pair] : pair), undecided
//               pairs to und_idx
// The fast decisions equal the exact ones: the two amplitudes differ by less than eps (gate_eps, plus the
// truncation bound: dropping taps changes the exact blurred value by at most norm * (ksum^2 - ksum_t^2) times
// the largest |grid value|, itself at most the side's bound b).
template <int MODE, int DEC>
__global__ __launch_bounds__(128) void k_amp(const SideParams *sides, const uint32_t *sel, uint32_t n, GateKernel K,
                                             double thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp,
                                             uint32_t *count, uint32_t cap, float *all_amp, const uint32_t *n_dev,
                                             uint32_t *und_idx, uint32_t *und_count) {
    __shared__ WaveLds L[2];
    __shared__ float red[2], bnd[2];
    constexpr bool Q1 = MODE == 1 && DEC == GATE_SPLIT_Q;   // the 16-bit tables of the 512^2 grid
    __shared__ uint32_t TQ[Q1 ? 2 : 1][Q1 ? 768 : 1];
    const uint32_t item = blockIdx.x;
    if (n_dev) n = min(n, *n_dev);
    if (item >= n) return;
    const uint32_t pidx = sel ? sel[item] : item;
    const int wv = threadIdx.x >> 6;
    float b = 0.f;
    constexpr bool TR = DEC == GATE_SPLIT_T || (DEC == GATE_SPLIT_Q && MODE == 0);   // truncated blur
    const float m = side_max<MODE, DEC != GATE_EXACT, TR, DEC == GATE_SPLIT_Q>(sides[2 * size_t(pidx) + wv], L[wv], K, &b,
                                                                               Q1 ? TQ[Q1 ? wv : 0] : nullptr);
    if (DEC == GATE_SPLIT || DEC == GATE_SPLIT_T) b = side_bound<MODE>(L[wv]);
    if ((threadIdx.x & 63) == 0) {
        red[wv] = m;
        bnd[wv] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float amp = 0;
        amp += red[0];
        amp += red[1];
        if (DEC == GATE_EXACT) {
            if (all_amp) all_amp[item] = amp;
            if (double(amp) >= thr) {
                const uint32_t i = atomicAdd(count, 1u);
                if (i < cap) {
                    out_idx[i] = kmap ? kmap[pidx] : pidx;
                    if (out_amp) out_amp[i] = amp;
                }
            }
        } else {
            // a truncated blur moves each side's maximum by at most K.trunc times its grid bound; the
            // fixed-point grid by at most 2.83e-4 times it, and the integer blur passes (their rounded weights,
            // the horizontal sums rounded to 16 bits) by at most K.wq times it
            const double bs = double(bnd[0]) + double(bnd[1]);
            const double eps = gate_eps(bnd[0], bnd[1]) + (TR ? double(K.trunc) * bs : 0.0) +
                               (DEC == GATE_SPLIT_Q ? ((MODE == 0 ? 2.9e-4 : 1.3e-4) + double(K.wq)) * bs : 0.0);
            if (all_amp) {   // (tests) the fast amplitude and its margin, by block
                all_amp[item] = amp;
                if (out_amp) out_amp[item] = float(eps);
            }
            if (double(amp) - eps >= thr) {
                const uint32_t i = atomicAdd(count, 1u);
                if (i < cap) out_idx[i] = kmap ? kmap[pidx] : pidx;
            } else if (double(amp) + eps >= thr) {
                const uint32_t i = atomicAdd(und_count, 1u);
                if (i < cap) und_idx[i] = pidx;
            }
        }
    }
}

// The reference's arithmetic for the pairs sel[0 .. min(*n_dev, grid)) in bands: work item w = (pair, side,
// band), one 64-thread workgroup (one wave) each, grid-stride over the items. A band's output rows take their
// values from the same rows of the horizontal pass as the whole grid's do, so each value, and a side's
// maximum over its bands, is the whole-grid kernel's; bandmax[w] gets the band's maximum (unnormalised).
template <int MODE> struct GateBands { static constexpr int NB = Shape<MODE>::G / 64; };
template <int MODE>
__global__ __launch_bounds__(64) void k_amp_band(const SideParams *sides, const uint32_t *sel, GateKernel K,
                                                 const uint32_t *n_dev, uint32_t grid, float *bandmax) {
    __shared__ WaveLds L;
    constexpr int NB = GateBands<MODE>::NB, NC = Shape<MODE>::G / 64, R = GateR<MODE>::R;
    const uint32_t n = min(*n_dev, grid);
    for (uint32_t w = blockIdx.x; w < n * 2u * NB; w += gridDim.x) {
        const uint32_t item = w / (2u * NB), side = (w / NB) & 1u, band = w % NB;
        const SideParams &sp = sides[2 * size_t(sel[item]) + side];
        __syncthreads();   // the previous item's reads of L are done
        grid_setup<MODE>(sp, L, K.wd);
        float best = -1e9f;
        grid_blur_rows<MODE, R, false, false>(sp, L, K.k, int(band) * 64, int(band) * 64 + 63, [&](int, const float (&o)[NC]) {
#pragma unroll
            for (int q = 0; q < NC; q++) best = fmaxf(best, o[q]);
        });
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o));
        if (threadIdx.x == 0) bandmax[w] = best;
    }
}

// Per pair of the banded list: each side's maximum over its bands times norm, the amplitude as k_amp sums it
// (0 + A + B), and k_amp's GATE_EXACT decision and output.
template <int MODE>
__global__ __launch_bounds__(256) void k_amp_band_decide(const uint32_t *sel, const uint32_t *n_dev, uint32_t grid,
                                                         const float *bandmax, float norm, double thr, const uint32_t *kmap,
                                                         uint32_t *out_idx, float *out_amp, uint32_t *count, uint32_t cap) {
    constexpr int NB = GateBands<MODE>::NB;
    const uint32_t n = min(*n_dev, grid);
    for (uint32_t item = blockIdx.x * 256u + threadIdx.x; item < n; item += gridDim.x * 256u) {
        float m[2];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            float best = -1e9f;
            for (int b = 0; b < NB; b++) best = fmaxf(best, bandmax[(item * 2u + side) * NB + b]);
            m[side] = best * norm;
        }
        float amp = 0;
        amp += m[0];
        amp += m[1];
        if (double(amp) >= thr) {
            const uint32_t pidx = sel[item];
            const uint32_t i = atomicAdd(count, 1u);
            if (i < cap) {
                out_idx[i] = kmap ? kmap[pidx] : pidx;
                if (out_amp) out_amp[i] = amp;
            }
        }
    }
}
constexpr uint32_t BAND_GRID = 2048, BAND_DECIDE_GRID = 16;

// ------------------------------------------------------------------ the gate as one chain without host round trips
__global__ void k_gate_clear(uint32_t *cnt) {
    if (threadIdx.x < GATE_NCNT) cnt[threadIdx.x] = 0;
}

// The number of phase survivors the later stages take: all, or none if they overflow the chain's grid.
__global__ void k_gate_clamp(uint32_t *cnt, uint32_t grid) {
    if (threadIdx.x == 0) cnt[3] = cnt[0] <= grid ? cnt[0] : 0u;
}

// One thread per full-gate survivor: its rank among the survivors' (unique) seed indices is its output slot.
__global__ __launch_bounds__(256) void k_gate_sort(const uint32_t *cnt, const uint32_t *k3, const float *a3, uint64_t start,
                                                   uint32_t grid, GateOut O) {
    const uint32_t n = cnt[3] ? min(cnt[2], grid) : 0u;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == 0) {
        O.counts[0] = cnt[0];
        O.counts[1] = cnt[3] ? cnt[1] : 0u;
        O.counts[2] = n;
        O.counts[3] = uint64_t(n) < O.cap ? uint64_t(n) : O.cap;
    }
    if (i >= n) return;
    const uint32_t key = k3[i];
    uint32_t rank = 0;
    for (uint32_t j = 0; j < n; j++) rank += k3[j] < key ? 1u : 0u;
    if (rank >= O.cap) return;
    if (O.seeds) O.seeds[rank] = mix64(start + key);
    if (O.amps) O.amps[rank] = a3[i];
    if (O.kidx) O.kidx[rank] = key;
}

// The amplitude launches of a chain: pre-gate over the phase survivors, full gate over its survivors. In the
// fast chain (C.fast) the pre-gate decides with the fast amplitude (GATE_SPLIT) and the exact pre-gate
// decides the undecided pairs (listed in k3, which only the full gate writes later). The full gate always
// runs the reference's arithmetic: it outputs the amplitudes, and a fast filter before it (measured: 15 %
// faster than the exact kernel, whose re-run of the ~16 % kept seeds then costs more) does not pay.
std::vector<GateLaunch> gate_launches(const GateChain &C, const GateStages &S) {
    const int mp = S.mode_pre;
    std::vector<GateLaunch> v;
    if (C.fast) {
        // the truncated fast amplitude where its wider margin stays small (the production kernel: 5.4e-4 per
        // unit of grid bound); DSX_GATE_TRUNC=0 keeps every tap
        static const bool trunc_on = [] {
            const char *e = std::getenv("DSX_GATE_TRUNC");
            return !(e && *e == '0');
        }();
        // and the grid in 16-bit fixed point (2.9e-4 more); DSX_GATE_Q16=0 keeps the float grid sums
        static const bool q16_on = [] {
            const char *e = std::getenv("DSX_GATE_Q16");
            return !(e && *e == '0');
        }();
        const bool tr = trunc_on && mp == 0 && S.kpre.trunc > 0.f && S.kpre.trunc <= 2e-3f;
        const int dec = tr ? (q16_on ? GATE_SPLIT_Q : GATE_SPLIT_T) : GATE_SPLIT;
        // (the pre-gate's undecided pairs, some 2 % of the phase survivors, keep a CU each busy for a 256^2
        // grid just as long as in bands, which measured 0.44 against 0.42 ms per 2^24 indices)
        v.push_back({mp, dec, true, nullptr, nullptr, C.i2, nullptr, C.cnt + 1, C.cnt + 3, C.k3, C.cnt + 4});
        v.push_back({mp, GATE_EXACT, true, C.k3, nullptr, C.i2, nullptr, C.cnt + 1, C.cnt + 4, nullptr, nullptr});
    } else {
        v.push_back({mp, GATE_EXACT, true, nullptr, nullptr, C.i2, nullptr, C.cnt + 1, C.cnt + 3, nullptr, nullptr});
    }
    if (C.fast && S.full_fast && C.u2) {
        // survivors only: the full gate decided from its 16-bit fast amplitude, the reference's arithmetic
        // for the undecided pairs (the amplitudes of the seeds decided fast are not computed), in bands: about
        // 1 % of the pre-gate survivors, whose 512^2 grids would each keep one CU busy for 0.6 ms
        v.push_back({1, GATE_SPLIT_Q, false, C.i2, C.k1, C.k3, nullptr, C.cnt + 2, C.cnt + 1, C.u2, C.cnt + 5});
        v.push_back({1, C.band ? GATE_EXACT_BANDS : GATE_EXACT, false, C.u2, C.k1, C.k3, C.a3, C.cnt + 2, C.cnt + 5,
                     nullptr, nullptr});
    } else {
        v.push_back({1, GATE_EXACT, false, C.i2, C.k1, C.k3, C.a3, C.cnt + 2, C.cnt + 1, nullptr, nullptr});
    }
    return v;
}

const void *amp_kernel(int mode, int dec) {
    if (dec == GATE_SPLIT_T && mode == 0) return reinterpret_cast<const void *>(&k_amp<0, GATE_SPLIT_T>);
    if (dec == GATE_SPLIT_Q && mode == 0) return reinterpret_cast<const void *>(&k_amp<0, GATE_SPLIT_Q>);
    if (dec == GATE_SPLIT_Q && mode == 1) return reinterpret_cast<const void *>(&k_amp<1, GATE_SPLIT_Q>);
    if (dec != GATE_EXACT)
        return mode == 0 ? reinterpret_cast<const void *>(&k_amp<0, GATE_SPLIT>)
                         : reinterpret_cast<const void *>(&k_amp<2, GATE_SPLIT>);
    return mode == 0 ? reinterpret_cast<const void *>(&k_amp<0, GATE_EXACT>)
           : mode == 1 ? reinterpret_cast<const void *>(&k_amp<1, GATE_EXACT>)
                       : reinterpret_cast<const void *>(&k_amp<2, GATE_EXACT>);
}

hipError_t enqueue_gate_chain(hipStream_t st, const GateChain &C, const GateStages &S, uint64_t start, uint32_t count,
                              const GateOut &O) {
    hipLaunchKernelGGL(k_gate_clear, dim3(1), dim3(64), 0, st, C.cnt);
    if (count) {
        launch_phase(st, start, count, S.maxph, C.k1, C.cnt, C.grid);
        hipLaunchKernelGGL(k_gate_clamp, dim3(1), dim3(64), 0, st, C.cnt, C.grid);
        launch_init(st, start, C.k1, nullptr, C.grid, C.sides, C.cnt + 3);
        for (const GateLaunch &g : gate_launches(C, S)) {
            const GateKernel &K = g.pre ? S.kpre : S.kfull;
            const double thr = g.pre ? S.thr_pre : S.amp;
            if (g.dec == GATE_EXACT_BANDS) {
                if (g.mode == 0) {
                    hipLaunchKernelGGL(k_amp_band<0>, dim3(BAND_GRID), dim3(64), 0, st, C.sides, g.sel, K, g.ndev, C.grid, C.band);
                    hipLaunchKernelGGL(k_amp_band_decide<0>, dim3(BAND_DECIDE_GRID), dim3(256), 0, st, g.sel, g.ndev, C.grid,
                                       C.band, K.norm, thr, g.kmap, g.out, g.amp, g.count, C.grid);
                } else {
                    hipLaunchKernelGGL(k_amp_band<1>, dim3(BAND_GRID), dim3(64), 0, st, C.sides, g.sel, K, g.ndev, C.grid, C.band);
                    hipLaunchKernelGGL(k_amp_band_decide<1>, dim3(BAND_DECIDE_GRID), dim3(256), 0, st, g.sel, g.ndev, C.grid,
                                       C.band, K.norm, thr, g.kmap, g.out, g.amp, g.count, C.grid);
                }
                continue;
            }
            launch_amp(st, g.mode, g.dec, C.sides, g.sel, C.grid, K, thr, g.kmap, g.out, g.amp, g.count, C.grid, nullptr,
                       g.ndev, g.und, g.und_count);
        }
    }
    hipLaunchKernelGGL(k_gate_sort, dim3((C.grid + 255) / 256), dim3(256), 0, st, C.cnt, C.k3, C.a3, start, C.grid, O);
    return hipGetLastError();
}

namespace {
hipKernelNodeParams kparams(const void *func, dim3 grid, dim3 block, void **args) {
    hipKernelNodeParams p{};
    p.func = const_cast<void *>(func);
    p.gridDim = grid;
    p.blockDim = block;
    p.kernelParams = args;
    p.sharedMemBytes = 0;
    p.extra = nullptr;
    return p;
}
// The argument lists of the nodes whose arguments change per chunk (start, count).
hipKernelNodeParams phase_params(const GateGraph &G, uint64_t &base, uint32_t &n, double &maxph, uint32_t *&out,
                                 uint32_t *&cnt, uint32_t &cap, void **a) {
    a[0] = &base; a[1] = &n; a[2] = &maxph; a[3] = &out; a[4] = &cnt; a[5] = &cap;
    return kparams(reinterpret_cast<const void *>(&k_phase), dim3((G.nmax + PHASE_PER_WG - 1) / PHASE_PER_WG), dim3(256), a);
}
}  // namespace

hipError_t gate_graph_build(GateGraph &G, const GateChain &C, const GateStages &S, uint32_t nmax, const GateOut &O) {
    gate_graph_free(G);
    G.C = C;
    G.S = S;
    G.O = O;
    G.nmax = nmax;
    hipError_t e = hipGraphCreate(&G.graph, 0);
    if (e != hipSuccess) return e;
    hipGraphNode_t prev = nullptr, node = nullptr;
    auto add = [&](const hipKernelNodeParams &p, hipGraphNode_t *out) -> hipError_t {
        const hipError_t r = hipGraphAddKernelNode(&node, G.graph, prev ? &prev : nullptr, prev ? 1 : 0, &p);
        if (r == hipSuccess) {
            prev = node;
            if (out) *out = node;
        }
        return r;
    };
    uint32_t *cnt = G.C.cnt;
    {   // clear
        void *a[] = {&cnt};
        if ((e = add(kparams(reinterpret_cast<const void *>(&k_gate_clear), dim3(1), dim3(64), a), nullptr))) return e;
    }
    {   // phase filter (grid for nmax indices; the kernel checks k < n)
        uint64_t base = 0;
        uint32_t n = nmax, cap = G.C.grid;
        double maxph = S.maxph;
        uint32_t *out = G.C.k1;
        void *a[6];
        if ((e = add(phase_params(G, base, n, maxph, out, cnt, cap, a), &G.n_phase))) return e;
    }
    {   // clamp
        uint32_t grid = G.C.grid;
        void *a[] = {&cnt, &grid};
        if ((e = add(kparams(reinterpret_cast<const void *>(&k_gate_clamp), dim3(1), dim3(64), a), nullptr))) return e;
    }
    {   // Perlin init of both sides of each phase survivor
        uint64_t base = 0;
        const uint32_t *ks = G.C.k1;
        const uint64_t *seeds = nullptr;
        uint32_t n = G.C.grid;
        SideParams *out = G.C.sides;
        const uint32_t *ndev = cnt + 3;
        void *a[] = {&base, &ks, &seeds, &n, &out, &ndev};
        if ((e = add(kparams(reinterpret_cast<const void *>(&k_init), dim3((2 * G.C.grid + 63) / 64), dim3(64), a),
                     &G.n_init)))
            return e;
    }
    for (const GateLaunch &g : gate_launches(G.C, S)) {   // pre-gate, then full gate
        const SideParams *sides = G.C.sides;
        const uint32_t *sel = g.sel, *kmap = g.kmap, *ndev = g.ndev;
        uint32_t n = G.C.grid, cap = G.C.grid;
        GateKernel K = g.pre ? S.kpre : S.kfull;
        double thr = g.pre ? S.thr_pre : S.amp;
        uint32_t *oidx = g.out, *count = g.count, *und = g.und, *und_count = g.und_count;
        float *oamp = g.amp, *all = nullptr;
        if (g.dec == GATE_EXACT_BANDS) {
            float *band = G.C.band;
            const float *cband = G.C.band;
            float norm = K.norm;
            void *a1[] = {&sides, &sel, &K, &ndev, &n, &band};
            void *a2[] = {&sel, &ndev, &n, &cband, &norm, &thr, &kmap, &oidx, &oamp, &count, &cap};
            const void *f1 = g.mode == 0 ? reinterpret_cast<const void *>(&k_amp_band<0>)
                                         : reinterpret_cast<const void *>(&k_amp_band<1>);
            const void *f2 = g.mode == 0 ? reinterpret_cast<const void *>(&k_amp_band_decide<0>)
                                         : reinterpret_cast<const void *>(&k_amp_band_decide<1>);
            if ((e = add(kparams(f1, dim3(BAND_GRID), dim3(64), a1), nullptr))) return e;
            if ((e = add(kparams(f2, dim3(BAND_DECIDE_GRID), dim3(256), a2), nullptr))) return e;
            continue;
        }
        void *a[] = {&sides, &sel, &n, &K, &thr, &kmap, &oidx, &oamp, &count, &cap, &all, &ndev, &und, &und_count};
        if ((e = add(kparams(amp_kernel(g.mode, g.dec), dim3(G.C.grid), dim3(128), a), nullptr))) return e;
    }
    {   // rank sort into the outputs
        const uint32_t *cc = cnt, *k3 = G.C.k3;
        const float *a3 = G.C.a3;
        uint64_t start = 0;
        uint32_t grid = G.C.grid;
        GateOut o = G.O;
        void *a[] = {&cc, &k3, &a3, &start, &grid, &o};
        if ((e = add(kparams(reinterpret_cast<const void *>(&k_gate_sort), dim3((G.C.grid + 255) / 256), dim3(256), a),
                     &G.n_sort)))
            return e;
    }
    return hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0);
}

hipError_t gate_graph_launch(GateGraph &G, hipStream_t st, uint64_t start, uint32_t count) {
    if (!G.exec || count > G.nmax) return hipErrorInvalidValue;
    hipError_t e;
    {
        uint64_t base = start;
        uint32_t n = count, cap = G.C.grid;
        double maxph = G.S.maxph;
        uint32_t *out = G.C.k1, *cnt = G.C.cnt;
        void *a[6];
        const hipKernelNodeParams p = phase_params(G, base, n, maxph, out, cnt, cap, a);
        if ((e = hipGraphExecKernelNodeSetParams(G.exec, G.n_phase, &p))) return e;
    }
    {
        uint64_t base = start;
        const uint32_t *ks = G.C.k1;
        const uint64_t *seeds = nullptr;
        uint32_t n = G.C.grid;
        SideParams *out = G.C.sides;
        const uint32_t *ndev = G.C.cnt + 3;
        void *a[] = {&base, &ks, &seeds, &n, &out, &ndev};
        const hipKernelNodeParams p =
            kparams(reinterpret_cast<const void *>(&k_init), dim3((2 * G.C.grid + 63) / 64), dim3(64), a);
        if ((e = hipGraphExecKernelNodeSetParams(G.exec, G.n_init, &p))) return e;
    }
    {
        const uint32_t *cc = G.C.cnt, *k3 = G.C.k3;
        const float *a3 = G.C.a3;
        uint64_t s0 = start;
        uint32_t grid = G.C.grid;
        GateOut o = G.O;
        void *a[] = {&cc, &k3, &a3, &s0, &grid, &o};
        const hipKernelNodeParams p =
            kparams(reinterpret_cast<const void *>(&k_gate_sort), dim3((G.C.grid + 255) / 256), dim3(256), a);
        if ((e = hipGraphExecKernelNodeSetParams(G.exec, G.n_sort, &p))) return e;
    }
    return hipGraphLaunch(G.exec, st);
}

void gate_graph_free(GateGraph &G) {
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G.exec = nullptr;
    G.graph = nullptr;
    G.n_phase = G.n_init = G.n_sort = nullptr;
}

// ------------------------------------------------------------------ launchers
void launch_phase(hipStream_t st, uint64_t base, uint32_t n, double maxph, uint32_t *out, uint32_t *count, uint32_t cap) {
    hipLaunchKernelGGL(k_phase, dim3((n + PHASE_PER_WG - 1) / PHASE_PER_WG), dim3(256), 0, st, base, n, maxph, out, count, cap);
}
void launch_init(hipStream_t st, uint64_t base, const uint32_t *ks, const uint64_t *seeds, uint32_t n, SideParams *out,
                 const uint32_t *n_dev) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_init, dim3((2 * n + 63) / 64), dim3(64), 0, st, base, ks, seeds, n, out, n_dev);
}
void launch_amp(hipStream_t st, int mode, int dec, const SideParams *sides, const uint32_t *sel, uint32_t n,
                const GateKernel &K, double thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp, uint32_t *count,
                uint32_t cap, float *all_amp, const uint32_t *n_dev, uint32_t *und_idx, uint32_t *und_count) {
    if (n == 0) return;
    void *a[] = {&sides, &sel, &n, const_cast<GateKernel *>(&K), &thr, &kmap, &out_idx, &out_amp, &count, &cap, &all_amp,
                 &n_dev, &und_idx, &und_count};
    (void)hipLaunchKernel(amp_kernel(mode, dec), dim3(n), dim3(128), a, 0, st);
}

}  // namespace dsx
