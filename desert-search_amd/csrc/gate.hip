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

namespace dsx {

// ------------------------------------------------------------------ y-phase filter
__global__ __launch_bounds__(256) void k_phase(uint64_t base, uint32_t n, double maxph, uint32_t *out, uint32_t *count,
                                               uint32_t cap) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    bool pass = false;
    if (k < n) pass = temperature_phase(mix64(base + k), nullptr) <= maxph;
    const uint64_t mask = __ballot(pass);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    uint32_t wbase = 0;
    if (lane == 0) wbase = atomicAdd(count, uint32_t(__popcll(mask)));
    wbase = __shfl(wbase, 0);
    if (pass) {
        const uint32_t idx = wbase + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
        if (idx < cap) out[idx] = k;
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

template <int MODE>
__device__ float side_max(const SideParams &sp, WaveLds &L, const GateKernel &K) {
    grid_setup<MODE>(sp, L, K.wd);
    float best = -1e9f;
    constexpr int NC = Shape<MODE>::G / 64;
    grid_blur_rows<MODE, GateR<MODE>::R>(sp, L, K.k, 0, Shape<MODE>::G - 1, [&](int, const float (&o)[NC]) {
#pragma unroll
        for (int q = 0; q < NC; q++) best = fmaxf(best, o[q]);
    });
    // wave maximum; x -> fl(x * norm) is monotone, so max(s) * norm == max(s * norm)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o));
    return best * K.norm;
}

// One 128-thread block per seed: wave 0 = side A, wave 1 = side B. sel (optional) maps block -> side
// pair index. Passing seeds are appended as (kmap ? kmap[pair] : pair, amplitude); all_amp (optional)
// gets every amplitude, indexed by block. n_dev (optional): the number of seeds is min(n, *n_dev), read
// on the device (the grid covers n).
template <int MODE>
__global__ __launch_bounds__(128) void k_amp(const SideParams *sides, const uint32_t *sel, uint32_t n, GateKernel K,
                                             double thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp,
                                             uint32_t *count, uint32_t cap, float *all_amp, const uint32_t *n_dev) {
    __shared__ WaveLds L[2];
    __shared__ float red[2];
    const uint32_t item = blockIdx.x;
    if (n_dev) n = min(n, *n_dev);
    if (item >= n) return;
    const uint32_t pidx = sel ? sel[item] : item;
    const int wv = threadIdx.x >> 6;
    const float m = side_max<MODE>(sides[2 * size_t(pidx) + wv], L[wv], K);
    if ((threadIdx.x & 63) == 0) red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float amp = 0;
        amp += red[0];
        amp += red[1];
        if (all_amp) all_amp[item] = amp;
        if (double(amp) >= thr) {
            const uint32_t i = atomicAdd(count, 1u);
            if (i < cap) {
                out_idx[i] = kmap ? kmap[pidx] : pidx;
                if (out_amp) out_amp[i] = amp;
            }
        }
    }
}

// ------------------------------------------------------------------ launchers
void launch_phase(hipStream_t st, uint64_t base, uint32_t n, double maxph, uint32_t *out, uint32_t *count, uint32_t cap) {
    hipLaunchKernelGGL(k_phase, dim3((n + 255) / 256), dim3(256), 0, st, base, n, maxph, out, count, cap);
}
void launch_init(hipStream_t st, uint64_t base, const uint32_t *ks, const uint64_t *seeds, uint32_t n, SideParams *out,
                 const uint32_t *n_dev) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_init, dim3((2 * n + 63) / 64), dim3(64), 0, st, base, ks, seeds, n, out, n_dev);
}
void launch_amp(hipStream_t st, int mode, const SideParams *sides, const uint32_t *sel, uint32_t n, const GateKernel &K,
                double thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp, uint32_t *count, uint32_t cap,
                float *all_amp, const uint32_t *n_dev) {
    if (n == 0) return;
    if (mode == 0)
        hipLaunchKernelGGL(k_amp<0>, dim3(n), dim3(128), 0, st, sides, sel, n, K, thr, kmap, out_idx, out_amp, count, cap,
                           all_amp, n_dev);
    else if (mode == 1)
        hipLaunchKernelGGL(k_amp<1>, dim3(n), dim3(128), 0, st, sides, sel, n, K, thr, kmap, out_idx, out_amp, count, cap,
                           all_amp, n_dev);
    else
        hipLaunchKernelGGL(k_amp<2>, dim3(n), dim3(128), 0, st, sides, sel, n, K, thr, kmap, out_idx, out_amp, count, cap,
                           all_amp, n_dev);
}

}  // namespace dsx
