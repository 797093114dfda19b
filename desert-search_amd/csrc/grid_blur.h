// One wave computes a periodic octave grid of one noise side, blurs it separably and hands each
// output row to an epilogue. Shared by the seed gate (gate.hip) and the period tables (tables.hip).
//
// Grid modes (reference arithmetic, summation orders kept exactly):
//   MODE 0  256^2, each cell the mean of the 4 half-lattice parities through a merged 3x3 corner
//           table, 18 lookups per cell                          gate.cuh:218-235, 256-281
//   MODE 1  512^2 half-lattice grid, 8 corner lookups per cell  gate.cuh:237-244, 262-291;
//           identical to lowfield.hpp:14-55 tab_setup/tabulate_octave for the same octave weight
//   MODE 2  256^2 even-parity grid, 8 corner lookups per cell   gate.cuh:237-244 (G = 256)
// Blur: H[j][i] = sum_d k[d] V[j][i+d-R], then O[j][i] = sum_d k[d] H[j+d-R][i], both periodic,
// accumulated from 0 in tap order (gate.cuh:296-313, lowfield.hpp:113-138, gproxy.cu:153-179).
//
// Layout: each lane owns NC = G/64 adjacent columns and walks the rows in pairs; the corner tables
// (16-float sub-tables, one per corner term, indexed by gradient hash) and the packed hash streams
// (Q4[m] = byte offsets 4*(P[(m+t)&255] & 15), t = 0..3) sit in LDS; the horizontal halo comes from
// the neighbouring lanes by shuffle (lane 63 <-> lane 0 is the periodic wrap); the vertical window of
// 2R+1 rows stays in registers.
#pragma once
#include "dsx_internal.h"

namespace dsx {

template <int MODE> struct Shape;
template <> struct Shape<0> { static constexpr int G = 256, NSTREAM = 12; };
template <> struct Shape<1> { static constexpr int G = 512, NSTREAM = 20; };
template <> struct Shape<2> { static constexpr int G = 256, NSTREAM = 10; };

struct WaveLds {
    uint8_t P[256];
    uint8_t col[2][256];
    uint32_t Q4[256];
    float tab[32 * 16];
};

__device__ __forceinline__ float tab_at(const float *tab, int sub, uint32_t off) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(tab) + sub * 64 + off);
}
__device__ __forceinline__ uint32_t byte_of(uint32_t dw, int b) { return (dw >> (8 * b)) & 0xffu; }

// Fills L (permutation, hashed columns, Q4, corner tables) for one side. wd: octave weight (double).
// Ends with a block barrier: every wave of the block must call it.
template <int MODE>
__device__ void grid_setup(const SideParams &sp, WaveLds &L, double wd) {
    const int lane = threadIdx.x & 63;
    reinterpret_cast<uint32_t *>(L.P)[lane] = reinterpret_cast<const uint32_t *>(sp.perm)[lane];
    __syncthreads();
    const int iy = sp.iy;
    for (int c = lane; c < 256; c += 64) {
        const int pc = L.P[c];
        L.col[0][c] = L.P[(pc + iy) & 255];
        L.col[1][c] = L.P[(pc + iy + 1) & 255];
        uint32_t q = 0;
        for (int t = 0; t < 4; t++) q |= uint32_t((L.P[(c + t) & 255] & 15) * 4) << (8 * t);
        L.Q4[c] = q;
    }
    const double ty = sp.vy;
    if (MODE == 0) {
        double fr[2][2];
        int sh[2][2];
        for (int ax = 0; ax < 2; ax++) {
            const double o = ax ? sp.oz : sp.ox;
            for (int q = 0; q < 2; q++) {
                const double b = o + 0.5 * q;
                fr[ax][q] = b - floor(b);
                sh[ax][q] = int(floor(b) - floor(o));
            }
        }
        for (int e = lane; e < 288; e += 64) {
            const int h = e & 15, dy = (e >> 4) & 1, uv = e >> 5, u = uv / 3, v = uv % 3;
            double acc = 0;
            for (int px = 0; px < 2; px++)
                for (int pz = 0; pz < 2; pz++) {
                    const int dx = u - sh[0][px], dz = v - sh[1][pz];
                    if (dx < 0 || dx > 1 || dz < 0 || dz > 1) continue;
                    const double fx = fr[0][px], fz = fr[1][pz], tx = fade(fx), tz = fade(fz);
                    const double wc = (dx ? tx : 1 - tx) * (dy ? ty : 1 - ty) * (dz ? tz : 1 - tz);
                    acc += wc * grad(unsigned(h), fx - dx, sp.fy - dy, fz - dz);
                }
            L.tab[e] = float(0.25 * wd * acc);
        }
    } else {
        constexpr bool HALF = (Shape<MODE>::G == 512);
        constexpr int NP = HALF ? 2 : 1;
        for (int e = lane; e < NP * NP * 128; e += 64) {
            const int par = e / 128, c = (e / 16) & 7, h = e & 15, px = par / 2, pz = par & 1;
            const double bx = (HALF ? px * 0.5 : 0.0) + sp.ox, bz = (HALF ? pz * 0.5 : 0.0) + sp.oz;
            const double fx = bx - floor(bx), fz = bz - floor(bz), tx = fade(fx), tz = fade(fz);
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const double wc = (dx ? tx : 1 - tx) * (dy ? ty : 1 - ty) * (dz ? tz : 1 - tz);
            L.tab[e] = float(wd * wc * grad(unsigned(h), fx - dx, sp.fy - dy, fz - dz));
        }
    }
    __syncthreads();
}

// Walks output rows [o_lo, o_hi] (o_hi - o_lo < G); epi(o, s[NC]) receives each blurred row segment
// of this lane (columns NC*lane .. NC*lane+NC-1), unnormalised.
template <int MODE, int R, class Epi>
__device__ void grid_blur_rows(const SideParams &sp, const WaveLds &L, const float *kern, int o_lo, int o_hi, Epi &&epi) {
    using S = Shape<MODE>;
    constexpr int G = S::G, NC = G / 64, W = 2 * R + 1;
    static_assert(R <= NC, "halo must come from the adjacent lane only");
    const int lane = threadIdx.x & 63;
    constexpr bool HALFO = (G == 512) || (MODE == 0);
    int ix0[2], iz0[2];
    for (int par = 0; par < 2; par++) {
        ix0[par] = int(floor((HALFO ? par * 0.5 : 0.0) + sp.ox));
        iz0[par] = int(floor((HALFO ? par * 0.5 : 0.0) + sp.oz));
    }
    int sbase[S::NSTREAM];
    if (MODE == 0) {
        for (int x = 0; x < 6; x++)
            for (int dy = 0; dy < 2; dy++) sbase[x * 2 + dy] = int(L.col[dy][(4 * lane + x + ix0[0]) & 255]) + iz0[0];
    } else {
        constexpr int NP = (G == 512) ? 2 : 1;
        for (int px = 0; px < NP; px++)
            for (int m = 0; m < 5; m++)
                for (int dy = 0; dy < 2; dy++)
                    sbase[(px * 5 + m) * 2 + dy] = int(L.col[dy][(4 * lane + ix0[px] + m) & 255]) + iz0[0];
    }
    const uint32_t diz = uint32_t(iz0[1] - iz0[0]);
    float kw[W];
#pragma unroll
    for (int d = 0; d < W; d++) kw[d] = kern[d];
    float ring[W][NC];
    const int js = (o_lo - R) & ~1;                        // first row (even)
    const int npair = (o_hi + R - js + 2) / 2;
    const float *tab = L.tab;
#pragma unroll 1
    for (int pr = 0; pr < npair; pr++) {
        const int j0 = js + 2 * pr;
        const int zoff = (MODE == 1) ? (j0 >> 1) : j0;
        uint32_t dw[S::NSTREAM];
#pragma unroll
        for (int s = 0; s < S::NSTREAM; s++) dw[s] = L.Q4[(sbase[s] + zoff) & 255];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int j = j0 + t;
            float V[NC];
            if (MODE == 0) {
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    float acc = 0;
#pragma unroll
                    for (int v = 0; v < 3; v++)
#pragma unroll
                        for (int dy = 0; dy < 2; dy++)
#pragma unroll
                            for (int u = 0; u < 3; u++)
                                acc += tab_at(tab, (u * 3 + v) * 2 + dy, byte_of(dw[(q + u) * 2 + dy], t + v));
                    V[q] = acc;
                }
            } else if (MODE == 1) {
                const int pz = t;
                const uint32_t zb = pz ? diz : 0u;
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    const int px = q & 1, mq = q >> 1;
                    float acc = 0;
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                        const uint32_t d = dw[(px * 5 + mq + dx) * 2 + dy];
                        const uint32_t off = (d >> (8u * (uint32_t(dz) + zb))) & 0xffu;
                        acc += tab_at(tab, (px * 2 + pz) * 8 + c, off);
                    }
                    V[q] = acc;
                }
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    float acc = 0;
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                        acc += tab_at(tab, c, byte_of(dw[(q + dx) * 2 + dy], t + dz));
                    }
                    V[q] = acc;
                }
            }
            float Wv[NC + 2 * R];
            const int left = (lane + 63) & 63, right = (lane + 1) & 63;
#pragma unroll
            for (int q = 0; q < R; q++) Wv[q] = __shfl(V[NC - R + q], left);
#pragma unroll
            for (int q = 0; q < NC; q++) Wv[R + q] = V[q];
#pragma unroll
            for (int q = 0; q < R; q++) Wv[R + NC + q] = __shfl(V[q], right);
#pragma unroll
            for (int k = 0; k < W - 1; k++)
#pragma unroll
                for (int q = 0; q < NC; q++) ring[k][q] = ring[k + 1][q];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                float s = 0;
#pragma unroll
                for (int d = 0; d < W; d++) s += kw[d] * Wv[q + d];
                ring[W - 1][q] = s;
            }
            const int o = j - R;
            if (o >= o_lo && o <= o_hi) {
                float out[NC];
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    float s = 0;
#pragma unroll
                    for (int d = 0; d < W; d++) s += kw[d] * ring[d][q];
                    out[q] = s;
                }
                epi(o, out);
            }
        }
    }
}

}  // namespace dsx
