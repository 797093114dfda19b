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
// (Q4[QPAD + m] = byte offsets 4*(P[(m+t)&255] & 15), t = 0..3, replicated for m in [-QPAD, QN - QPAD)
// so that a stream's row offset needs no wrap) sit in LDS; the horizontal halo comes from the
// neighbouring lanes by DPP wave rotates or LDS permutes (lane 63 <-> lane 0 is the periodic wrap); the vertical window of
// 2R+1 rows stays in registers. The corner tables come first in WaveLds: a table read is its hash
// byte's address plus an immediate sub-table offset, and paired reads (ds_read2, 8-bit dword offsets)
// reach 1 KB from one address.
#pragma once
#include "dsx_internal.h"

namespace dsx {

template <int MODE> struct Shape;
template <> struct Shape<0> { static constexpr int G = 256, NSTREAM = 12; };
template <> struct Shape<1> { static constexpr int G = 512, NSTREAM = 20; };
template <> struct Shape<2> { static constexpr int G = 256, NSTREAM = 10; };

// Hash-stream replication: a stream's base column is reduced to [0, 256) and the rows walked add offsets
// of at least -4 and at most 259 for every grid and blur radius used (checked in grid_blur_rows).
constexpr int QPAD = 8, QN = 528;

struct WaveLds {
    float tab[32 * 16];
    uint32_t Q4[QN];
    uint8_t P[256];
    uint8_t col[2][256];
};

// LDS byte address (offset in the workgroup's LDS) of a __shared__ object, and a 32-bit load from one
typedef __attribute__((address_space(3))) const uint32_t lds_u32;
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
    return uint32_t(reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) const void *)p));
}
__device__ __forceinline__ uint32_t lds_load(uint32_t addr) {
    return *reinterpret_cast<lds_u32 *>(uintptr_t(addr));
}

__device__ __forceinline__ float tab_at(const float *tab, int sub, uint32_t off) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(tab) + sub * 64 + off);
}
__device__ __forceinline__ uint32_t byte_of(uint32_t dw, int b) { return (dw >> (8 * b)) & 0xffu; }
// Two 16-bit integers in one register: sums of them are packed 16-bit adds (two cells per instruction)
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t tabq_u32(const uint32_t *tq, int sub, uint32_t off) {
    return *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(tq) + sub * 64 + off);
}
// The value of the lane to the left (lane - 1, lane 0 takes lane 63's) and to the right (lane + 1, lane 63
// takes lane 0's): DPP wavefront rotates (wave_ror:1, wave_rol:1), which use no LDS bandwidth
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x13C, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF, false));
}
// (every lane receives a value, so no old value needs setting: mov_dpp)
__device__ __forceinline__ s16x2 from_left(s16x2 v) {
    return __builtin_bit_cast(s16x2, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x13C, 0xF, 0xF, false));
}
__device__ __forceinline__ s16x2 from_right(s16x2 v) {
    return __builtin_bit_cast(s16x2, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x134, 0xF, 0xF, false));
}
// One step of a sum that starts from 0: s + x, or x itself for the sum's first term when ZSIGN is off
// (the two differ only when x is -0; see grid_blur_rows)
template <bool ZSIGN> __device__ __forceinline__ float sum_step(float s, float x, bool first) {
    return (!ZSIGN && first) ? x : s + x;
}

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
    }
    for (int c = lane; c < QN; c += 64) {
        uint32_t q = 0;
        for (int t = 0; t < 4; t++) q |= uint32_t((L.P[(c - QPAD + t) & 255] & 15) * 4) << (8 * t);
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

// The grid in 16-bit fixed point, for the fast amplitude only. The float sub-tables set up by grid_setup are
// scaled by S = 32000 / b (b: the side's grid bound, the sum of the largest |entry| of the sub-tables one
// cell reads) and rounded to integers, then packed so that one 32-bit read gives the terms of a pair of
// cells that read the same lattice point (low half: the pair's first cell, high half: its second, 0 where
// a cell reads no such term), into tq:
//   MODE 0 (18 sub-tables T(u, v, dy), 384 words in L.tab): adjacent cells (c, c + 1); table
//          (v * 2 + dy) * 4 + p is lattice column c + p: low round(S T(p, v, dy)) (p <= 2), high
//          round(S T(p - 1, v, dy)) (p >= 1)
//   MODE 1 (8 sub-tables T(par, dx + 2 dy + 4 dz) per parity par = px * 2 + pz, 768 words): cells (q, q + 2)
//          of one parity; table ((par * 2 + dy) * 2 + dz) * 3 + p is lattice column mq + p of the pair's
//          first cell: low round(S T(par, p, dy, dz)) (p <= 1), high round(S T(par, p - 1, dy, dz)) (p >= 1)
// Each entry is biased by its sub-table's largest |rounded entry| OFF, so that both halves hold values in
// [0, 2 OFF]: a cell's terms then add up exactly as 32-bit integers, both halves at once (three per
// v_add3_u32), with no carry from the low half into the high one: a half's sum lies in [0, 2 B], B = the sum
// of its cell's OFFs <= S b + 18 * 0.502, and 2 B < 65536. Subtracting B from each half on its own (a packed
// 16-bit subtraction; qbias: B in both halves, of MODE 0 or per parity of MODE 1) leaves the signed sum,
// |sum| <= B < 32767. Each term is within 0.502 of S times
// its float (the rounding to an integer, plus the float product's error of at most 32000 * 2^-24), so the
// integer sum of n terms is within 0.502 n of S times the exact sum: 2.83e-4 * b (MODE 0, 18 terms) or
// 1.26e-4 * b (MODE 1, 8 terms) after the division by S. Every wave of the block must call it (block
// barriers); returns S (1 when b is below 1e-20: then every term rounds to 0, off by less than b).
template <int MODE>
__device__ float grid_setup_q16(WaveLds &L, float b, uint32_t *tq, uint32_t (&qbias)[4]) {
    static_assert(MODE == 0 || MODE == 1, "16-bit grids of the 256^2 and 512^2 gates");
    const int lane = threadIdx.x & 63;
    const float S = b > 1e-20f ? 32000.f / b : 1.f;
    constexpr int NSUB = MODE == 0 ? 18 : 32, NW = MODE == 0 ? 6 : 12;   // float sub-tables; words per lane
    const auto q = [&](int sub, int h) { return int(rintf(L.tab[sub * 16 + h] * S)); };
    // lane i < NSUB: sub-table i's offset
    int off = 0;
    if (lane < NSUB)
        for (int h = 0; h < 16; h++) off = max(off, abs(q(lane, h)));
    // B: the sum over all 18 sub-tables (MODE 0) or over the 8 of each parity (MODE 1, lanes par * 8 ..)
    int bs = off;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) bs += __shfl_xor(bs, o);
    if (MODE == 0) {
#pragma unroll
        for (int o = 8; o < 64; o <<= 1) bs += __shfl_xor(bs, o);
    }
#pragma unroll
    for (int par = 0; par < 4; par++) qbias[par] = uint32_t(__shfl(bs, MODE == 0 ? 0 : par * 8)) * 0x10001u;
    uint32_t w[NW];
#pragma unroll
    for (int k = 0; k < NW; k++) {
        const int e = lane + 64 * k, sub = e >> 4, h = e & 15;
        int slo = -1, shi = -1;   // float sub-tables of the low and high half (-1: no term)
        if (MODE == 0) {
            const int vd = sub >> 2, p = sub & 3;
            if (p <= 2) slo = p * 6 + vd;
            if (p >= 1) shi = (p - 1) * 6 + vd;
        } else {
            const int p = sub % 3, pdd = sub / 3, dz = pdd & 1, dy = (pdd >> 1) & 1, par = pdd >> 2;
            if (p <= 1) slo = par * 8 + p + 2 * dy + 4 * dz;
            if (p >= 1) shi = par * 8 + (p - 1) + 2 * dy + 4 * dz;
        }
        const int olo = __shfl(off, slo < 0 ? 0 : slo), ohi = __shfl(off, shi < 0 ? 0 : shi);
        const int lo = slo >= 0 ? q(slo, h) + olo : 0, hi = shi >= 0 ? q(shi, h) + ohi : 0;
        w[k] = uint32_t(lo) | (uint32_t(hi) << 16);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NW; k++) tq[lane + 64 * k] = w[k];
    __syncthreads();
    return S;
}

// Walks output rows [o_lo, o_hi] (0 <= o_lo <= o_hi < G); epi(o, s[NC]) receives each blurred row
// segment of this lane (columns NC*lane .. NC*lane+NC-1), unnormalised.
// ZSIGN: every sum starts from 0 as the reference's loops do. Without it (for callers that keep only the
// maximum) a sum starts from its first term: 0 + x == x except that 0 + (-0) == +0, so values differ at
// most in the sign of a zero, and a maximum does not.
// FMA: the two blur passes accumulate with fused multiply-adds (one rounding per tap instead of two, 9 or
// 15 fewer instructions per pass and cell). Not the reference's arithmetic: each pass is then within
// gamma_W = W u / (1 - W u) (u = 2^-24) of the exact dot product times sum |k_d| |input|, as the
// unfused pass is, so the two results differ by at most 2 gamma_W of that per pass (see gate.hip).
// Q16 (with FMA; MODE 0 at R = 3, MODE 1 at R = 7, after grid_setup_q16 into tq): the grid values are the
// 16-bit fixed-point sums, in units of 1/S, and both blur passes run on them in integer arithmetic: 16-bit
// dot products (v_dot2_i32_i16) of two adjacent cells, or of two adjacent rows, with the weight pairs kq2
// (the W taps w times SK, rounded: SK = 2^14 at R = 3, 2^13 at R = 7; with NP = R + 1 pairs, kq2[m] =
// (w[2m - 1], w[2m]) for a pair's first cell or row and kq2[NP + m] = (w[2m], w[2m + 1]) for its second,
// w[-1] = w[W] = 0). The horizontal sums, biased by 2^15, are rounded to their high 16 bits between the
// passes, so epi receives int values in units of 2^16 / (SK^2 S).
template <int MODE, int R, bool ZSIGN = true, bool FMA = false, bool Q16 = false, class Epi>
__device__ void grid_blur_rows(const SideParams &sp, const WaveLds &L, const float *kern, int o_lo, int o_hi, Epi &&epi,
                               const uint32_t *kq2 = nullptr, const uint32_t *tq = nullptr,
                               const uint32_t *qbias = nullptr) {
    using S = Shape<MODE>;
    constexpr int G = S::G, NC = G / 64, W = 2 * R + 1;
    static_assert(R <= NC, "halo must come from the adjacent lane only");
    // rows j0 run from (o_lo - R) & ~1 >= (-R) & ~1 up to o_hi + R <= G - 1 + R; a stream reads row
    // offset zoff = j0 (j0 >> 1 for the half-lattice grid) past a base column in [0, 256)
    constexpr int ZMIN = (MODE == 1) ? (((-R) & ~1) >> 1) : ((-R) & ~1);
    constexpr int ZMAX = (MODE == 1) ? ((G - 1 + R) >> 1) : (G - 1 + R);
    static_assert(QPAD + ZMIN >= 0 && 255 + QPAD + ZMAX < QN, "stream rows beyond the replicated hash stream");
    const int lane = threadIdx.x & 63;
    constexpr bool HALFO = (G == 512) || (MODE == 0);
    int ix0[2], iz0[2];
    for (int par = 0; par < 2; par++) {
        ix0[par] = int(floor((HALFO ? par * 0.5 : 0.0) + sp.ox));
        iz0[par] = int(floor((HALFO ? par * 0.5 : 0.0) + sp.oz));
    }
    // per stream: the LDS address of its row-0 hash dword; row offset zoff adds 4 * zoff
    uint32_t sbase[S::NSTREAM];
    const auto q4_addr = [&](int c) { return lds_addr(&L.Q4[QPAD + ((c + iz0[0]) & 255)]); };
    if (MODE == 0) {
        for (int x = 0; x < 6; x++)
            for (int dy = 0; dy < 2; dy++) sbase[x * 2 + dy] = q4_addr(L.col[dy][(4 * lane + x + ix0[0]) & 255]);
    } else {
        constexpr int NP = (G == 512) ? 2 : 1;
        for (int px = 0; px < NP; px++)
            for (int m = 0; m < 5; m++)
                for (int dy = 0; dy < 2; dy++)
                    sbase[(px * 5 + m) * 2 + dy] = q4_addr(L.col[dy][(4 * lane + ix0[px] + m) & 255]);
    }
    const uint32_t diz = uint32_t(iz0[1] - iz0[0]);
    float kw[W];
#pragma unroll
    for (int d = 0; d < W; d++) kw[d] = kern[d];
    constexpr bool QH = Q16;
    static_assert(!Q16 || (FMA && NC == R + 1 && (MODE == 0 || MODE == 1)),
                  "the integer passes take NC / 2 = (R + 1) / 2 cell pairs per lane");
    constexpr int NP = R + 1;   // (QH) weight pairs per pass
    s16x2 kq[QH ? 2 * NP : 1];
    if constexpr (QH) {
#pragma unroll
        for (int m = 0; m < 2 * NP; m++) kq[m] = __builtin_bit_cast(s16x2, kq2[m]);
    }
    // The vertical window: with UNR (the fast 256^2 grid at radius 3 or less) row j sits in slot
    // (j - js) % WR of a ring of WR = W + 1 slots and the loop runs WR / 2 row pairs per iteration, so every
    // slot index is a constant and nothing moves; otherwise the W rows shift down one slot per row. (At
    // radius 4 the unrolled loop needs 153 VGPRs: 3 waves per SIMD instead of 4.)
    constexpr bool UNR = FMA && ((MODE == 0 && R <= 3) || QH);
    constexpr int WR = UNR ? W + 1 : W, PU = UNR ? WR / 2 : 1;
    static_assert(!QH || PU == NP, "the integer vertical pass keeps R + 1 row pairs");
    float ring[QH ? 1 : WR][NC];   // (float passes)
    // (QH) the vertical window as 16-bit row pairs: slot u holds rows (j0, j0 + 1) of pair u, each value
    // the horizontal sum rounded to units of 2^16 (its high half after adding 2^15)
    s16x2 hring[QH ? NP : 1][NC] = {};
    int hrow[2][NC];
    const int js = (o_lo - R) & ~1;                        // first row (even)
    const int npair = (o_hi + R - js + 2) / 2;
    const float *tab = L.tab;
#pragma unroll 1
    for (int pr0 = 0; pr0 < npair; pr0 += PU)
#pragma unroll
    for (int u = 0; u < PU; u++) {
        const int pr = pr0 + u;
        if (UNR && pr >= npair) break;
        const int j0 = js + 2 * pr;
        const int zoff = (MODE == 1) ? (j0 >> 1) : j0;
        uint32_t dw[S::NSTREAM];
#pragma unroll
        for (int s = 0; s < S::NSTREAM; s++)
            dw[s] = lds_load(sbase[s] + uint32_t(4 * zoff));
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int j = j0 + t;
            if constexpr (QH) {
                // the row's cells as 16-bit adjacent pairs A[a] = (cell 2a, cell 2a + 1)
                s16x2 A[NC / 2];
                if constexpr (MODE == 0) {
                    // cells (2 pq, 2 pq + 1) together: lattice column 2 pq + p of each (v, dy) term
#pragma unroll
                    for (int pq = 0; pq < NC / 2; pq++) {
                        uint32_t acc = 0;
#pragma unroll
                        for (int v = 0; v < 3; v++)
#pragma unroll
                            for (int dy = 0; dy < 2; dy++)
#pragma unroll
                                for (int p = 0; p < 4; p++) {
                                    const uint32_t e = tabq_u32(tq, (v * 2 + dy) * 4 + p, byte_of(dw[(2 * pq + p) * 2 + dy], t + v));
                                    acc = (v == 0 && dy == 0 && p == 0) ? e : acc + e;
                                }
                        A[pq] = __builtin_bit_cast(s16x2, acc) - __builtin_bit_cast(s16x2, qbias[0]);   // per half
                    }
                } else {
                    // cells (px + 4k, px + 4k + 2) of parity (px, t) together: lattice column 2k + p of each
                    // (dy, dz) term; row t reads hash byte dz + t * diz (the stream shifted once per row)
                    s16x2 B[2][2];
#pragma unroll
                    for (int px = 0; px < 2; px++)
#pragma unroll
                        for (int k = 0; k < 2; k++) {
                            uint32_t acc = 0;
#pragma unroll
                            for (int dy = 0; dy < 2; dy++)
#pragma unroll
                                for (int dz = 0; dz < 2; dz++)
#pragma unroll
                                    for (int p = 0; p < 3; p++) {
                                        const uint32_t d = dw[(px * 5 + 2 * k + p) * 2 + dy];
                                        const uint32_t ds = t ? (d >> (8u * diz)) : d;
                                        const uint32_t e = tabq_u32(tq, (((px * 2 + t) * 2 + dy) * 2 + dz) * 3 + p, byte_of(ds, dz));
                                        acc = (dy == 0 && dz == 0 && p == 0) ? e : acc + e;
                                    }
                            B[px][k] = __builtin_bit_cast(s16x2, acc) - __builtin_bit_cast(s16x2, qbias[px * 2 + t]);
                        }
                    // (q, q + 2) pairs of both parities into adjacent pairs: (4k, 4k + 1), (4k + 2, 4k + 3)
#pragma unroll
                    for (int k = 0; k < 2; k++) {
                        const uint32_t b0 = __builtin_bit_cast(uint32_t, B[0][k]), b1 = __builtin_bit_cast(uint32_t, B[1][k]);
                        A[2 * k] = __builtin_bit_cast(s16x2, __builtin_amdgcn_perm(b1, b0, 0x05040100u));
                        A[2 * k + 1] = __builtin_bit_cast(s16x2, __builtin_amdgcn_perm(b1, b0, 0x07060302u));
                    }
                }
                // horizontal pass: pairs X = the left lane's NC / 2, this lane's, the right lane's; cell c's W
                // taps are the NP pairs from X[(c + 1) / 2] with the weight pairs of its parity
                s16x2 X[3 * (NC / 2)];
#pragma unroll
                for (int a = 0; a < NC / 2; a++) {
                    X[a] = from_left(A[a]);
                    X[NC / 2 + a] = A[a];
                    X[NC + a] = from_right(A[a]);
                }
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    int h = 1 << 15;
#pragma unroll
                    for (int m = 0; m < NP; m++) h = __builtin_amdgcn_sdot2(X[(c + 1) / 2 + m], kq[(c & 1) * NP + m], h, false);
                    hrow[t][c] = h;
                }
            } else {
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
                                    acc = sum_step<ZSIGN>(acc, tab_at(tab, (u * 3 + v) * 2 + dy, byte_of(dw[(q + u) * 2 + dy], t + v)),
                                                          v == 0 && dy == 0 && u == 0);
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
                            acc = sum_step<ZSIGN>(acc, tab_at(tab, (px * 2 + pz) * 8 + c, off), c == 0);
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
                            acc = sum_step<ZSIGN>(acc, tab_at(tab, c, byte_of(dw[(q + dx) * 2 + dy], t + dz)), c == 0);
                        }
                        V[q] = acc;
                    }
                }
                // the halo: DPP rotates for the 256^2 grids (their kernels wait on LDS issue: measured -6 %), LDS
                // permutes for the 512^2 grid (its 2 x 7 halo values per row as rotates measured +8 %)
                float Wv[NC + 2 * R];
                const int left = (lane + 63) & 63, right = (lane + 1) & 63;
#pragma unroll
                for (int q = 0; q < R; q++) Wv[q] = G == 256 ? from_left(V[NC - R + q]) : __shfl(V[NC - R + q], left);
#pragma unroll
                for (int q = 0; q < NC; q++) Wv[R + q] = V[q];
#pragma unroll
                for (int q = 0; q < R; q++) Wv[R + NC + q] = G == 256 ? from_right(V[q]) : __shfl(V[q], right);
                if constexpr (!UNR) {
#pragma unroll
                    for (int k = 0; k < W - 1; k++)
#pragma unroll
                        for (int q = 0; q < NC; q++) ring[k][q] = ring[k + 1][q];
                }
                // slot of row j - W + 1 + d, the d-th row of the window ending at row j (d = W - 1: row j)
                const auto slot = [&](int d) { return UNR ? (2 * u + t - (W - 1) + d + 2 * WR) % WR : d; };
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    float s = 0;
#pragma unroll
                    for (int d = 0; d < W; d++)
                        s = FMA ? (d == 0 ? kw[0] * Wv[q] : __builtin_fmaf(kw[d], Wv[q + d], s))
                                : sum_step<ZSIGN>(s, kw[d] * Wv[q + d], d == 0);
                    ring[slot(W - 1)][q] = s;
                }
                const int o = j - R;
                if (o >= o_lo && o <= o_hi) {
                    float out[NC];
#pragma unroll
                    for (int q = 0; q < NC; q++) {
                        float s = 0;
#pragma unroll
                        for (int d = 0; d < W; d++)
                            s = FMA ? (d == 0 ? kw[0] * ring[slot(0)][q] : __builtin_fmaf(kw[d], ring[slot(d)][q], s))
                                    : sum_step<ZSIGN>(s, kw[d] * ring[slot(d)][q], d == 0);
                        out[q] = s;
                    }
                    epi(o, out);
                }
            }
        }
        if constexpr (QH) {
            // vertical pass over row pairs u - R .. u (rows j0 - 2R .. j0 + 1): output j0 - R takes the odd
            // weight pairs, output j0 - R + 1 the even ones, as the horizontal pass does for a cell pair
#pragma unroll
            for (int c = 0; c < NC; c++)
                hring[u][c] = __builtin_bit_cast(s16x2, __builtin_amdgcn_perm(uint32_t(hrow[1][c]), uint32_t(hrow[0][c]), 0x07060302u));
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int o = j0 - R + e;
                if (o >= o_lo && o <= o_hi) {
                    int out[NC];
#pragma unroll
                    for (int c = 0; c < NC; c++) {
                        int v = 0;
#pragma unroll
                        for (int m = 0; m < NP; m++)
                            v = __builtin_amdgcn_sdot2(hring[(u + 1 + m) % NP][c], kq[(1 - e) * NP + m], v, false);
                        out[c] = v;
                    }
                    epi(o, out);
                }
            }
        }
    }
}

}  // namespace dsx
