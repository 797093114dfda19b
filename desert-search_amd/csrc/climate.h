// FP64 climate sampler for gfx950: improved Perlin noise, 2-sided normal noise, terrain spline,
// surface-biome grid and parameter-tree fallback, and the flood predicates built on them.
//
// Reference semantics, restated as device code with the same double/float operation order:
//   Perlin init / sample         standalone.hpp:55-82
//   NormalNoise init / sample    standalone.hpp:123-151
//   spline                       standalone.hpp:154-169
//   classify (nearest leaf)      standalone.hpp:170-196 (iterative here: explicit frame stack)
//   climate / finish             standalone.hpp:224-238
//   SurfaceGrid::lookup          surface.hpp:58-70 (grid built on the host, dsx_api)
//   desert / desert_state        surface.hpp:85-124
//   DesertProxy                  desert_proxy.hpp:14-32
//   hot predicate                dsearch_desert.cpp:334
// A seed's 46 octaves (T 2, H 2, C 9, E 4, Shift 3, W 3 per side) are initialised once per seed on
// the device (k_climinit) and read from LDS by the flood kernels.
#pragma once
#include "dsx_internal.h"
#include <type_traits>
#include "../data/mc_tables.inc"

namespace dsx {

constexpr int CLIM_NOCT = 46;
// octave slots per channel and side, and the first slot of each channel (side 0, then side 1)
__host__ __device__ constexpr int clim_n(int ch) { return ch == 0 ? 2 : ch == 1 ? 2 : ch == 2 ? 9 : ch == 3 ? 4 : 3; }
__host__ __device__ constexpr int clim_base(int ch) {
    return ch == 0 ? 0 : ch == 1 ? 4 : ch == 2 ? 8 : ch == 3 ? 26 : ch == 4 ? 34 : 40;
}

// Minecraft's 16 gradient vectors (standalone.hpp:40-49) as doubles: (gx, gy) pairs and gz in two
// arrays, so that the hash-indexed 16-byte and 8-byte LDS reads of a wave hit distinct banks.
struct GradTab {
    double2 xy[16];
    double z[16];
};
struct ClimOct {
    uint8_t perm[256];
    double ox, oy, oz, freq, weight, fy, vy;
    int iy, pad_;
};
struct ClimSeed {
    GradTab g;
    ClimOct oct[CLIM_NOCT];
    double factor[6];
};

// Surface-biome grid (surface.hpp:12-71): cut positions per axis T,H,C,E,W and the cell table.
struct SurfGrid {
    const int32_t *cuts[5];
    int ncut[5];
    int dim[5];
    const int16_t *cell;
};

// ------------------------------------------------------------------ noise
__device__ __forceinline__ double lerpd(double t, double a, double b) { return a + t * (b - a); }

// grad() of dsx_rng.h with the components read from the table: the same products and sum order.
__device__ __forceinline__ double grad_t(const GradTab &G, int h, double x, double y, double z) {
    const double2 g = G.xy[h & 15];
    return g.x * x + g.y * y + G.z[h & 15] * z;
}

__device__ __forceinline__ double perlin_sample(const GradTab &G, const ClimOct &o, double x, double y, double z) {
    x += o.ox;
    z += o.oz;
    const double bx = floor(x), bz = floor(z);
    const int ix = int(bx), iz = int(bz);
    x -= bx;
    z -= bz;
    int jy = o.iy;
    double dy = o.fy, ty = o.vy;
    if (y != 0) {
        y += o.oy;
        const double by = floor(y);
        jy = int(by);
        dy = y - by;
        ty = fade(dy);
    }
    const uint8_t *P = o.perm;
    const int a0 = P[ix & 255], a1 = P[(ix + 1) & 255];
    const int b00 = P[(a0 + jy) & 255], b01 = P[(a0 + jy + 1) & 255];
    const int b10 = P[(a1 + jy) & 255], b11 = P[(a1 + jy + 1) & 255];
    const double tx = fade(x), tz = fade(z);
    const double x1 = x - 1, y1 = dy - 1, z1 = z - 1;
    const double l00 = lerpd(tx, grad_t(G, P[(b00 + iz) & 255], x, dy, z), grad_t(G, P[(b10 + iz) & 255], x1, dy, z));
    const double l10 = lerpd(tx, grad_t(G, P[(b01 + iz) & 255], x, y1, z), grad_t(G, P[(b11 + iz) & 255], x1, y1, z));
    const double l01 = lerpd(tx, grad_t(G, P[(b00 + iz + 1) & 255], x, dy, z1), grad_t(G, P[(b10 + iz + 1) & 255], x1, dy, z1));
    const double l11 = lerpd(tx, grad_t(G, P[(b01 + iz + 1) & 255], x, y1, z1), grad_t(G, P[(b11 + iz + 1) & 255], x1, y1, z1));
    return lerpd(tz, lerpd(ty, l00, l10), lerpd(ty, l01, l11));
}

// The temperature octaves alone (slots 0-3), enough for the hot predicate.
struct ClimT {
    GradTab g;
    ClimOct oct[4];
    double factor[6];
};

// The 14 octaves the proxy predicate reads (T both sides, C octaves 0-2 and E octaves 0-1 of both
// sides; desert_proxy.hpp:14-32): the slots of ClimSeed in this order.
constexpr int PROXY_SLOTS[14] = {0, 1, 2, 3, 8, 9, 10, 17, 18, 19, 26, 27, 30, 31};
struct ClimP {
    GradTab g;
    ClimOct oct[14];
    double factor[6];
};

static_assert(offsetof(ClimT, oct) == offsetof(ClimSeed, oct), "ClimT is a prefix of ClimSeed");
static_assert(offsetof(ClimP, oct) == offsetof(ClimSeed, oct), "ClimP starts like ClimSeed");
static_assert(sizeof(ClimSeed) % 16 == 0 && offsetof(ClimT, factor) % 16 == 0 && offsetof(ClimP, factor) % 16 == 0 &&
                  sizeof(ClimOct) % 16 == 0,
              "uint4 copies");

// Slot of octave k of `side` of channel ch in a climate block CS. ClimSeed and ClimT hold the slots in
// ClimSeed order; ClimP holds PROXY_SLOTS (only the octaves the proxy reads are valid).
template <class CS>
__host__ __device__ constexpr int oct_slot(int ch, int side, int k) {
    if constexpr (std::is_same<CS, ClimP>::value) {
        return ch == DSX_CH_T ? side * 2 + k : ch == DSX_CH_C ? 4 + side * 3 + k : 10 + side * 2 + k;
    } else {
        return clim_base(ch) + side * clim_n(ch) + k;
    }
}

// Copy the part of a seed's climate that CS holds (all of it, or ClimT's prefix and the factors);
// threads lane, lane + nt, ... of the block participate.
template <class CS>
__device__ __forceinline__ void copy_climate(CS &dst, const ClimSeed &src, int lane, int nt) {
    const uint4 *s = reinterpret_cast<const uint4 *>(&src);
    uint4 *d = reinterpret_cast<uint4 *>(&dst);
    if constexpr (std::is_same<CS, ClimP>::value) {
        constexpr int GW = int(sizeof(GradTab) / 16), OW = int(sizeof(ClimOct) / 16);
        for (int k = lane; k < GW + 14 * OW; k += nt) {
            const int o = (k - GW) / OW, w = (k - GW) % OW;
            d[k] = k < GW ? s[k] : s[GW + PROXY_SLOTS[o] * OW + w];
        }
    } else {
        constexpr int pre = int(offsetof(CS, factor) / 16);
        for (int k = lane; k < pre; k += nt) d[k] = s[k];
    }
    if (lane < 6) dst.factor[lane] = src.factor[lane];
}

// NormalNoise::sample: per-side sums, then (A + B) * factor. CS: ClimSeed, or ClimT for channel T.
template <class CS>
__device__ __forceinline__ double noise_sample(const CS &C, int ch, double x, double y, double z) {
    const int n = clim_n(ch);
    double sum0 = 0, sum1 = 0;
    {
        const double px = x, py = y, pz = z;
#pragma unroll 1
        for (int j = 0; j < n; j++) {
            const ClimOct &o = C.oct[oct_slot<CS>(ch, 0, j)];
            sum0 += o.weight * perlin_sample(C.g, o, px * o.freq, py * o.freq, pz * o.freq);
        }
    }
    {
        const double sc = 337. / 331.;
        const double px = x * sc, py = y * sc, pz = z * sc;
#pragma unroll 1
        for (int j = 0; j < n; j++) {
            const ClimOct &o = C.oct[oct_slot<CS>(ch, 1, j)];
            sum1 += o.weight * perlin_sample(C.g, o, px * o.freq, py * o.freq, pz * o.freq);
        }
    }
    return (sum0 + sum1) * C.factor[ch];
}
// DesertProxy::octaves: one running sum over both sides, octaves [0, hi).
template <class CS>
__device__ __forceinline__ double noise_octaves(const CS &C, int ch, double x, double z, int hi) {
    const int n = clim_n(ch);
    double s = 0;
    for (int side = 0; side < 2; side++) {
        const double sc = side ? 337.0 / 331.0 : 1.0;
#pragma unroll 1
        for (int k = 0; k < hi && k < n; k++) {
            const ClimOct &o = C.oct[oct_slot<CS>(ch, side, k)];
            s += o.weight * perlin_sample(C.g, o, x * sc * o.freq, 0, z * sc * o.freq);
        }
    }
    return s * C.factor[ch];
}

// ------------------------------------------------------------------ terrain spline (depth 4: unrolled by template)
template <int D>
__device__ __noinline__ float spline_at(int index, const float *v) {
    const dsx_spline_node &s = DSX_SPLINE[index];
    if (s.n == 0) return s.value;
    if constexpr (D == 0) {
        return s.value;   // unreachable: the spline tree is 4 levels deep
    } else {
        const float x = v[s.axis];
        int hi = 0;
        while (hi < s.n && x > s.x[hi]) hi++;
        if (hi == 0 || hi == s.n) {
            const int j = hi == 0 ? 0 : s.n - 1;
            return spline_at<D - 1>(s.child[j], v) + s.slope[j] * (x - s.x[j]);
        }
        const int lo = hi - 1;
        const float width = s.x[hi] - s.x[lo], t = (x - s.x[lo]) / width;
        const float a = spline_at<D - 1>(s.child[lo], v), b = spline_at<D - 1>(s.child[hi], v);
        const float p = s.slope[lo] * width - (b - a), q = -s.slope[hi] * width + (b - a);
        return float(lerpd(double(t), double(a), double(b)) + double(t * (1.f - t)) * lerpd(double(t), double(p), double(q)));
    }
}
__device__ __noinline__ float spline0(const float *v) { return spline_at<5>(0, v); }

// ------------------------------------------------------------------ biome classification
__device__ __forceinline__ uint64_t box_dist(int node, const int64_t *v) {
    const uint64_t enc = DSX_TREE[node];
    uint64_t sum = 0;
    for (int i = 0; i < 6; i++) {
        const int idx = int((enc >> (8 * i)) & 255);
        const int64_t lo = DSX_PARAM[idx][0], hi = DSX_PARAM[idx][1];
        int64_t d = lo - v[i];
        const int64_t d2 = v[i] - hi;
        if (d2 > d) d = d2;
        if (d < 0) d = 0;
        sum += uint64_t(d * d);
    }
    return sum;
}
// Nearest leaf by branch and bound, visiting children in the reference's order so that ties resolve
// to the same leaf (first found wins: strict < everywhere).
__device__ __noinline__ int classify(const int64_t *v) {
    struct Frame { int child, step, level, i; };
    Frame st[8];
    int sp = 0;
    uint64_t bestd = ~0ull;
    int best = 0;
    auto open = [&](int node, int level) {
        int step;
        do { step = DSX_STEPS[level++]; } while (node + step >= DSX_NTREE);
        st[sp++] = Frame{int(DSX_TREE[node] >> 48), step, level, 0};
    };
    if ((DSX_TREE[0] >> 56) == 255) {
        best = 0;
    } else {
        open(0, 0);
    }
    while (sp > 0) {
        Frame &f = st[sp - 1];
        if (f.i >= 6 || f.child >= DSX_NTREE) { sp--; continue; }
        const int c = f.child;
        f.child += f.step;
        f.i++;
        const uint64_t d = box_dist(c, v);
        if (!(d < bestd)) continue;
        if ((DSX_TREE[c] >> 56) == 255) {
            bestd = d;
            best = c;
        } else {
            open(c, f.level);
        }
    }
    return int((DSX_TREE[best] >> 48) & 255);
}

__device__ __forceinline__ int grid_lookup(const SurfGrid &g, const int64_t *q) {
    constexpr int AX[5] = {0, 1, 2, 3, 5};
    int idx[5];
#pragma unroll
    for (int a = 0; a < 5; a++) {
        const int64_t v = q[AX[a]];
        const int32_t *c = g.cuts[a];
        const int nc = g.ncut[a];
        if (v < c[0] || v >= c[nc - 1]) return -1;
        // upper_bound - 1
        int lo = 0, hi = nc;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (int64_t(c[mid]) <= v) lo = mid + 1;
            else hi = mid;
        }
        idx[a] = lo - 1;
        if (v == c[idx[a]]) return -1;
    }
    const size_t k = ((((size_t(idx[0]) * g.dim[1] + idx[1]) * g.dim[2] + idx[2]) * g.dim[3] + idx[3]) * g.dim[4] + idx[4]);
    return g.cell[k];
}

// ------------------------------------------------------------------ predicates
__device__ __forceinline__ float ridges(float w) { return -(fabsf(fabsf(w) - 0.6666667f) - 0.33333334f) * 3.f; }

__device__ __forceinline__ void shifted(const ClimSeed &C, int x, int z, double &px, double &pz) {
    px = x + 4. * noise_sample(C, DSX_CH_SHIFT, x, 0, z);
    pz = z + 4. * noise_sample(C, DSX_CH_SHIFT, z, x, 0);
}

// Sampler::climate at quart (x, y, z) (standalone.hpp:224-238).
__device__ void climate_q(const ClimSeed &C, int x, int y, int z, int64_t *q) {
    double px, pz;
    shifted(C, x, z, px, pz);
    const float c = float(noise_sample(C, DSX_CH_C, px, 0, pz));
    const float e = float(noise_sample(C, DSX_CH_E, px, 0, pz)), w = float(noise_sample(C, DSX_CH_W, px, 0, pz));
    const float v[4] = {c, e, ridges(w), w};
    const double off = double(spline0(v) + 0.015f);
    const float d = float(1. - (y * 4) / 128. - 83. / 160. + off);
    const float t = float(noise_sample(C, DSX_CH_T, px, 0, pz)), h = float(noise_sample(C, DSX_CH_H, px, 0, pz));
    const float vals[6] = {t, h, c, e, d, w};
    for (int i = 0; i < 6; i++) q[i] = int64_t(10000.f * vals[i]);
}

// surface.hpp:85-93
// *hot (when given) is set when the temperature test passes, i.e. when the 36 further octave samples
// of C, E, W and H are taken: the predicate's cost class.
__device__ bool is_desert(const ClimSeed &C, const SurfGrid &g, int x, int z, bool *hot = nullptr) {
    double px, pz;
    shifted(C, x, z, px, pz);
    const float t = float(noise_sample(C, DSX_CH_T, px, 0, pz));
    if (int64_t(10000.f * t) < 5500) return false;
    if (hot) *hot = true;
    const float c = float(noise_sample(C, DSX_CH_C, px, 0, pz));
    const float e = float(noise_sample(C, DSX_CH_E, px, 0, pz)), w = float(noise_sample(C, DSX_CH_W, px, 0, pz));
    const float v[4] = {c, e, ridges(w), w};
    const double off = double(spline0(v) + 0.015f);
    const float d = float(1. - (64 * 4) / 128. - 83. / 160. + off);
    const float h = float(noise_sample(C, DSX_CH_H, px, 0, pz));   // T is deterministic: reuse t
    int64_t q[6] = {int64_t(10000.f * t), int64_t(10000.f * h), int64_t(10000.f * c),
                    int64_t(10000.f * e),  int64_t(10000.f * d), int64_t(10000.f * w)};
    const int b = grid_lookup(g, q);
    return (b >= 0 ? b : classify(q)) == 2;
}

// surface.hpp:98-124: 2 desert, 1 river that would be desert with weirdness +-0.0501, 0 other.
__device__ int desert_state(const ClimSeed &C, const SurfGrid &g, int x, int z, bool *hot = nullptr) {
    double px, pz;
    shifted(C, x, z, px, pz);
    const float t = float(noise_sample(C, DSX_CH_T, px, 0, pz));
    if (int64_t(10000.f * t) < 5500) return 0;
    if (hot) *hot = true;
    const float c = float(noise_sample(C, DSX_CH_C, px, 0, pz)), e = float(noise_sample(C, DSX_CH_E, px, 0, pz));
    const float w = float(noise_sample(C, DSX_CH_W, px, 0, pz)), h = float(noise_sample(C, DSX_CH_H, px, 0, pz));
    int64_t q[6] = {int64_t(10000.f * t), int64_t(10000.f * h), int64_t(10000.f * c),
                    int64_t(10000.f * e), -10000,               int64_t(10000.f * w)};
    bool exact_depth = false;
    auto pick = [&]() {
        const int b = grid_lookup(g, q);
        if (b >= 0) return b;
        if (!exact_depth) {
            const float v[4] = {c, e, ridges(w), w};
            const double off = double(spline0(v) + 0.015f);
            q[4] = int64_t(10000.f * float(1. - (64 * 4) / 128. - 83. / 160. + off));
            exact_depth = true;
        }
        return classify(q);
    };
    const int b = pick();
    if (b == 2) return 2;
    if (b != 7 && b != 11) return 0;
    if (q[5] > 500 || q[5] < -500) return 0;
    q[5] = q[5] >= 0 ? 501 : -501;
    return pick() == 2 ? 1 : 0;
}

// The two sides of a noise computed apart (by two waves), in the operations of noise_sample and
// noise_octaves: side 0 and side 1 of noise_sample's sums at y = 0 (the caller forms (sum0 + sum1) *
// factor); side 0's running sum of noise_octaves from 0, or side 1's weighted terms, which the caller adds
// to side 0's sum in order before multiplying by factor.
template <class CS>
__device__ __forceinline__ double noise_side(const CS &C, int ch, int side, double x, double z) {
    const int n = clim_n(ch);
    const double sc = 337. / 331.;
    const double px = side ? x * sc : x, py = side ? 0. * sc : 0., pz = side ? z * sc : z;
    double sum = 0;
#pragma unroll 1
    for (int j = 0; j < n; j++) {
        const ClimOct &o = C.oct[oct_slot<CS>(ch, side, j)];
        sum += o.weight * perlin_sample(C.g, o, px * o.freq, py * o.freq, pz * o.freq);
    }
    return sum;
}
template <class CS>
__device__ __forceinline__ double octaves_side(const CS &C, int ch, int side, double x, double z, int hi, double *terms) {
    const int n = clim_n(ch);
    const double sc = side ? 337.0 / 331.0 : 1.0;
    double s = 0;
#pragma unroll 1
    for (int k = 0; k < hi && k < n; k++) {
        const ClimOct &o = C.oct[oct_slot<CS>(ch, side, k)];
        const double t = o.weight * perlin_sample(C.g, o, x * sc * o.freq, 0, z * sc * o.freq);
        if (side) terms[k] = t;
        else s += t;
    }
    return s;
}

// dsearch_desert.cpp:334
template <class CS>
__device__ __forceinline__ int hot_state(const CS &C, int x, int z) {
    return noise_sample(C, DSX_CH_T, x, 0, z) >= 0.55 ? 2 : 0;
}
// desert_proxy.hpp:25-32
template <class CS>
__device__ __forceinline__ float proxy_prob(const CS &C, int x, int z, bool *hot = nullptr) {
    if (noise_sample(C, DSX_CH_T, x, 0, z) < 0.55) return 0;
    if (hot) *hot = true;
    const double c = noise_octaves(C, DSX_CH_C, x, z, 3), e = noise_octaves(C, DSX_CH_E, x, z, 2);
    int i = int(floor((c + 0.8) / 0.05)), j = int(floor((e + 1.0) / 0.05));
    i = i < 0 ? 0 : (i > 39 ? 39 : i);
    j = j < 0 ? 0 : (j > 39 ? 39 : j);
    return DSX_DPROB[i][j];
}
template <class CS>
__device__ __forceinline__ int proxy_state(const CS &C, int x, int z, float cut, bool *hot = nullptr) {
    return proxy_prob(C, x, z, hot) >= cut ? 2 : 0;
}

enum PredKind { PRED_DESERT_STATE = 0, PRED_HOT = 1, PRED_PROXY = 2, PRED_DESERT = 3 };

__device__ __forceinline__ int predicate(int kind, const ClimSeed &C, const SurfGrid &g, int x, int z, float cut,
                                         bool *hot = nullptr) {
    switch (kind) {
    case PRED_DESERT_STATE: return desert_state(C, g, x, z, hot);
    case PRED_HOT: return hot_state(C, x, z);
    case PRED_PROXY: return proxy_state(C, x, z, cut, hot);
    default: return is_desert(C, g, x, z, hot) ? 2 : 0;
    }
}

// Octave samples (Perlin::sample calls) one predicate evaluation takes, by kind and cost class
// (cold: the temperature test failed). Shift 2 x 6, T 4, C 18, E 8, W 6, H 4 (standalone.hpp:94-100);
// the proxy reads C octaves 0-2 and E octaves 0-1 of both sides (desert_proxy.hpp:14-24).
__host__ __device__ constexpr int pred_samples(int kind, bool hot) {
    return kind == PRED_HOT ? 4 : kind == PRED_PROXY ? (hot ? 14 : 4) : (hot ? 52 : 16);
}
// FP64 operations of one octave sample as Perlin::sample writes them (standalone.hpp:62-82) with
// NormalNoise's weighting (standalone.hpp:146-149), y = 0: 3 coordinate scalings, 2 offsets, 2 floors,
// 2 fraction subtractions, 2 fades x 7, 3 corner offsets, 8 gradient dots x 5, 7 lerps x 3, weight
// and accumulate 2.
constexpr int FLOPS_PER_SAMPLE = 89;

// ------------------------------------------------------------------ FP32 pre-pass
// The noise in FP32 where that is safe to decide a predicate, FP64 (above) for the rest. The lattice
// cell and the fractions are computed in FP64 exactly as perlin_sample; fades, gradient dot products
// and lerps run in FP32 with FMA. Error bound (u = 2^-24), per octave sample:
//   fractions f in [0, 1]: |f32 - f| <= u;  fade: |fade'| <= 1.875, evaluation <= 80u  => <= 82u;
//   gradient dots (components -1, 0, 1, one of them 0; |terms| <= 1): <= 6u;
//   lerp a + t(b - a) with |b - a| <= 4: error <= max(e_a, e_b) + 4 e_t + 6u, so x-lerps <= 340u,
//   y-lerps <= 350u, z-lerp <= 684u (< 4.1e-5).
// A noise value sums its octaves (|weight| sum <= 2 for T) times factor <= 5/3, plus the FP32 sums:
//   |T32 - T64| <= 5/3 * 2 * 684u + 20u < 1.4e-4   =>  PRE_ERR_T = 2e-4 is used.
// At shifted coordinates (desert_state) the shift noise in FP32 moves px by at most 4 x 3e-4 quarts;
// improved noise has |grad| <= 3.75 * 2.45 + 1 < 10.2 per lattice unit and T's octaves scale quarts by
// at most 2^-9 * 337/331, so T moves by at most 10.2 * 1.2e-3 * 2^-9 * 1.02 * 2 * 5/3 < 1e-4 more:
// PRE_ERR_TS = 4e-4 is used. A cell is decided in FP32 only when T32 is farther than the bound from
// the threshold; every other cell is evaluated in FP64, so results are bit-identical to plain FP64.
constexpr double PRE_ERR_T = 2e-4, PRE_ERR_TS = 4e-4;

__device__ __forceinline__ float fade32(float t) { return t * t * t * fmaf(t, fmaf(t, 6.f, -15.f), 10.f); }
__device__ __forceinline__ float lerp32(float t, float a, float b) { return fmaf(t, b - a, a); }
__device__ __forceinline__ float grad32(const GradTab &G, int h, float x, float y, float z) {
    const double2 g = G.xy[h & 15];
    return fmaf(float(G.z[h & 15]), z, fmaf(float(g.y), y, float(g.x) * x));
}
__device__ __forceinline__ float perlin_sample32(const GradTab &G, const ClimOct &o, double x, double y, double z) {
    x += o.ox;
    z += o.oz;
    const double bx = floor(x), bz = floor(z);
    const int ix = int(bx), iz = int(bz);
    const float fx = float(x - bx), fz = float(z - bz);
    int jy = o.iy;
    float dy = float(o.fy), ty = float(o.vy);
    if (y != 0) {
        y += o.oy;
        const double by = floor(y);
        jy = int(by);
        dy = float(y - by);
        ty = fade32(dy);
    }
    const uint8_t *P = o.perm;
    const int a0 = P[ix & 255], a1 = P[(ix + 1) & 255];
    const int b00 = P[(a0 + jy) & 255], b01 = P[(a0 + jy + 1) & 255];
    const int b10 = P[(a1 + jy) & 255], b11 = P[(a1 + jy + 1) & 255];
    const float tx = fade32(fx), tz = fade32(fz);
    const float x1 = fx - 1.f, y1 = dy - 1.f, z1 = fz - 1.f;
    const float l00 = lerp32(tx, grad32(G, P[(b00 + iz) & 255], fx, dy, fz), grad32(G, P[(b10 + iz) & 255], x1, dy, fz));
    const float l10 = lerp32(tx, grad32(G, P[(b01 + iz) & 255], fx, y1, fz), grad32(G, P[(b11 + iz) & 255], x1, y1, fz));
    const float l01 = lerp32(tx, grad32(G, P[(b00 + iz + 1) & 255], fx, dy, z1), grad32(G, P[(b10 + iz + 1) & 255], x1, dy, z1));
    const float l11 = lerp32(tx, grad32(G, P[(b01 + iz + 1) & 255], fx, y1, z1), grad32(G, P[(b11 + iz + 1) & 255], x1, y1, z1));
    return lerp32(tz, lerp32(ty, l00, l10), lerp32(ty, l01, l11));
}
// NormalNoise::sample with FP32 octaves (coordinates scaled in FP64 as noise_sample does)
template <class CS>
__device__ __forceinline__ float noise_sample32(const CS &C, int ch, double x, double y, double z) {
    const int n = clim_n(ch);
    float sum0 = 0, sum1 = 0;
#pragma unroll 1
    for (int j = 0; j < n; j++) {
        const ClimOct &o = C.oct[oct_slot<CS>(ch, 0, j)];
        sum0 = fmaf(float(o.weight), perlin_sample32(C.g, o, x * o.freq, y * o.freq, z * o.freq), sum0);
    }
    const double sc = 337. / 331.;
    const double px = x * sc, py = y * sc, pz = z * sc;
#pragma unroll 1
    for (int j = 0; j < n; j++) {
        const ClimOct &o = C.oct[oct_slot<CS>(ch, 1, j)];
        sum1 = fmaf(float(o.weight), perlin_sample32(C.g, o, px * o.freq, py * o.freq, pz * o.freq), sum1);
    }
    return (sum0 + sum1) * float(C.factor[ch]);
}

// The predicates with the FP32 pre-pass on their temperature test: the hot test decided both ways,
// the proxy and the biome predicates' cold cells rejected; the rest in FP64.
template <class CS>
__device__ __forceinline__ int hot_state_pre(const CS &C, int x, int z) {
    const double t = noise_sample32(C, DSX_CH_T, x, 0, z);
    if (t + PRE_ERR_T < 0.55) return 0;
    if (t - PRE_ERR_T >= 0.55) return 2;
    return hot_state(C, x, z);
}
template <class CS>
__device__ __forceinline__ int proxy_state_pre(const CS &C, int x, int z, float cut, bool *hot = nullptr) {
    if (double(noise_sample32(C, DSX_CH_T, x, 0, z)) + PRE_ERR_T < 0.55) return 0;
    return proxy_state(C, x, z, cut, hot);
}
// desert_state / desert: 0 when T at the FP32-shifted position is certainly below 0.5499 (then
// int64(10000f * float(T)) <= 5499 < 5500, surface.hpp:102-105)
__device__ __forceinline__ bool biome_cold_pre(const ClimSeed &C, int x, int z) {
    const double px = x + 4. * double(noise_sample32(C, DSX_CH_SHIFT, x, 0, z));
    const double pz = z + 4. * double(noise_sample32(C, DSX_CH_SHIFT, z, x, 0));
    return double(noise_sample32(C, DSX_CH_T, px, 0, pz)) + PRE_ERR_TS < 0.5499;
}
__device__ __forceinline__ int predicate_pre(int kind, const ClimSeed &C, const SurfGrid &g, int x, int z, float cut,
                                             bool *hot = nullptr) {
    switch (kind) {
    case PRED_HOT: return hot_state_pre(C, x, z);
    case PRED_PROXY: return proxy_state_pre(C, x, z, cut, hot);
    default: return biome_cold_pre(C, x, z) ? 0 : predicate(kind, C, g, x, z, cut, hot);
    }
}
// the predicate of a flood or point batch: plain FP64 (pre = 0) or with the FP32 pre-pass (pre = 1)
__device__ __forceinline__ int predicate_mode(int pre, int kind, const ClimSeed &C, const SurfGrid &g, int x, int z,
                                              float cut, bool *hot = nullptr) {
    return pre ? predicate_pre(kind, C, g, x, z, cut, hot) : predicate(kind, C, g, x, z, cut, hot);
}

void launch_climinit(hipStream_t st, const uint64_t *seeds, const uint32_t *slots, uint32_t n, ClimSeed *out);

}  // namespace dsx
