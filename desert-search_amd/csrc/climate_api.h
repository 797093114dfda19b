// Launchers of the climate kernels (climate.hip), used by the API translation units.
#pragma once
#include "climate.h"

namespace dsx {
void launch_noise(hipStream_t st, const ClimSeed *cs, int ch, const double *pts, uint32_t n, double *out);
void launch_climate(hipStream_t st, const ClimSeed *cs, const int32_t *xs, const int32_t *zs, uint32_t n, int64_t *q);
void launch_pred(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, int kind, float cut, const int32_t *xs,
                 const int32_t *zs, uint32_t n, int8_t *out);
void launch_points(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, const uint32_t *sidx, const int32_t *xs,
                   const int32_t *zs, uint32_t n, int what, float cut, int8_t *out8, double *outd);
}  // namespace dsx
