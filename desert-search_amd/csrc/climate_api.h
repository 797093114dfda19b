// Launchers of the climate kernels (climate.hip), used by the API translation units.
#pragma once
#include "climate.h"

namespace dsx {
// fp32: the FP32 octaves of the pre-pass (noise_sample32) instead of FP64
void launch_noise(hipStream_t st, const ClimSeed *cs, int ch, const double *pts, uint32_t n, double *out, int fp32 = 0);
void launch_climate(hipStream_t st, const ClimSeed *cs, const int32_t *xs, const int32_t *zs, uint32_t n, int64_t *q);
// pre: 1 = the FP32 pre-pass on the temperature test (identical results)
void launch_pred(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, int kind, float cut, const int32_t *xs,
                 const int32_t *zs, uint32_t n, int8_t *out, int pre);
void launch_points(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, const uint32_t *sidx, const int32_t *xs,
                   const int32_t *zs, uint32_t n, int what, float cut, int8_t *out8, double *outd, int pre);
// proxy-table Monte Carlo (tools/dfunc.cpp): hot flags of tries, then the histogram of the selected hot points
void launch_dfunc_hot(hipStream_t st, const ClimSeed *cs, const uint32_t *sidx, const int32_t *xs, const int32_t *zs,
                      uint32_t n, uint8_t *hot);
void launch_dfunc_rec(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, const uint32_t *sidx, const int32_t *xs,
                      const int32_t *zs, uint32_t n, unsigned long long *hist);
}  // namespace dsx
