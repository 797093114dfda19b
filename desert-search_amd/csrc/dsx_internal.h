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
struct GateKernel {
    float k[15];
    float norm;
    double wd;   // octave weight the corner tables are scaled by
};

void launch_phase(hipStream_t st, uint64_t base, uint32_t n, double maxph, uint32_t *out, uint32_t *count, uint32_t cap);
// n_dev (optional): a device word holding the item count, read by the kernels (min(n, *n_dev)); the grid
// still covers n items, so a caller that does not know the count on the host launches with an upper bound.
void launch_init(hipStream_t st, uint64_t base, const uint32_t *ks, const uint64_t *seeds, uint32_t n, SideParams *out,
                 const uint32_t *n_dev = nullptr);
void launch_amp(hipStream_t st, int mode, const SideParams *sides, const uint32_t *sel, uint32_t n, const GateKernel &K,
                double thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp, uint32_t *count, uint32_t cap,
                float *all_amp, const uint32_t *n_dev = nullptr);

}  // namespace dsx
