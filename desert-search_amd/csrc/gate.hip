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
#include "dsx_internal.h"

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

__global__ __launch_bounds__(64) void k_init(uint64_t base, const uint32_t *ks, const uint64_t *seeds, uint32_t n,
                                             SideParams *out) {
    __shared__ uint32_t slot32[64 * PERM_STRIDE / 4];
    uint8_t *slots = reinterpret_cast<uint8_t *>(slot32);
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
// MODE 0: anti-aliased 256^2 pre-gate (4-parity mean, merged 3x3 corner table, R=4)  gate.cuh:218-235,256-281
// MODE 1: full 512^2 amplitude (all parities, 8 corners, R=7)                          gate.cuh:237-244,262-291
// MODE 2: legacy 256^2 even-parity pre-gate (8 corners, R=4)
template <int MODE> struct Shape;
template <> struct Shape<0> { static constexpr int G = 256, R = 4, NSUB = 18, NSTREAM = 12; };
template <> struct Shape<1> { static constexpr int G = 512, R = 7, NSUB = 32, NSTREAM = 20; };
template <> struct Shape<2> { static constexpr int G = 256, R = 4, NSUB = 8, NSTREAM = 10; };

// Per-wave LDS: permutation, hashed column rows, packed hash-offset stream table, corner tables.
struct WaveLds {
    uint8_t P[256];
    uint8_t col[2][256];
    uint32_t Q4[256];       // Q4[m] = bytes { 4*(P[(m+t)&255] & 15) : t = 0..3 }
    float tab[32 * 16];     // NSUB sub-tables of 16 floats (one per gradient hash)
    float red;
};

__device__ __forceinline__ float tab_at(const float *tab, int sub, uint32_t off) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(tab) + sub * 64 + off);
}
__device__ __forceinline__ uint32_t byte_of(uint32_t dw, int b) { return (dw >> (8 * b)) & 0xffu; }

template <int MODE>
__device__ float side_max(const SideParams &sp, WaveLds &L, const GateKernel &K) {
    using S = Shape<MODE>;
    constexpr int G = S::G, R = S::R, NC = G / 64, W = 2 * R + 1;
    const int lane = threadIdx.x & 63;
    // ---- setup (lane-parallel): P, col, Q4, corner tables
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
    const float w = K.w;
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
            L.tab[e] = float(0.25 * w * acc);
        }
    } else {
        constexpr bool HALF = (G == 512);
        constexpr int NP = HALF ? 2 : 1;
        for (int e = lane; e < NP * NP * 128; e += 64) {
            const int par = e / 128, c = (e / 16) & 7, h = e & 15, px = par / 2, pz = par & 1;
            const double bx = (HALF ? px * 0.5 : 0.0) + sp.ox, bz = (HALF ? pz * 0.5 : 0.0) + sp.oz;
            const double fx = bx - floor(bx), fz = bz - floor(bz), tx = fade(fx), tz = fade(fz);
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const double wc = (dx ? tx : 1 - tx) * (dy ? ty : 1 - ty) * (dz ? tz : 1 - tz);
            L.tab[e] = float(1.0 * w * wc * grad(unsigned(h), fx - dx, sp.fy - dy, fz - dz));
        }
    }
    __syncthreads();

    // lattice origins per parity (gate.cuh:249-253)
    constexpr bool HALFO = (G == 512) || (MODE == 0);
    int ix0[2], iz0[2];
    for (int par = 0; par < 2; par++) {
        ix0[par] = int(floor((HALFO ? par * 0.5 : 0.0) + sp.ox));
        iz0[par] = int(floor((HALFO ? par * 0.5 : 0.0) + sp.oz));
    }
    // ---- hash streams: base offsets into Q4 per (lattice column, dy)
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
    const uint32_t diz = uint32_t(iz0[1] - iz0[0]);   // 0 or 1 (only used by MODE 1)

    float ring[W][NC];
    float best = -1e9f;
    constexpr int JS = (-R) & ~1;                      // first row, even, <= -R
    constexpr int NPAIR = (G + R - JS + 1) / 2;        // row pairs covering rows JS .. G-1+R
    const float *tab = L.tab;
#pragma unroll 1
    for (int pr = 0; pr < NPAIR; pr++) {
        const int j0 = JS + 2 * pr;
        // one Q4 dword per stream covers both rows of the pair (z offsets 0..3)
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
                const int pz = t;                                  // j0 even
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
            // horizontal blur: halo columns from the neighbouring lanes (periodic across the wave)
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
                for (int d = 0; d < W; d++) s += K.k[d] * Wv[q + d];
                ring[W - 1][q] = s;
            }
            // vertical blur of output row j - R once rows j-2R .. j are in the window
            if (j >= R && j <= G - 1 + R) {
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    float s = 0;
#pragma unroll
                    for (int d = 0; d < W; d++) s += K.k[d] * ring[d][q];
                    best = fmaxf(best, s);
                }
            }
        }
    }
    // wave maximum; x -> fl(x * norm) is monotone, so max(s) * norm == max(s * norm)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o));
    return best * K.norm;
}

// One 128-thread block per seed: wave 0 = side A, wave 1 = side B. sel (optional) maps block -> side
// pair index. Passing seeds are appended as (kmap ? kmap[pair] : pair, amplitude); all_amp (optional)
// gets every amplitude, indexed by block.
template <int MODE>
__global__ __launch_bounds__(128) void k_amp(const SideParams *sides, const uint32_t *sel, uint32_t n, GateKernel K,
                                             float thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp,
                                             uint32_t *count, uint32_t cap, float *all_amp) {
    __shared__ WaveLds L[2];
    __shared__ float red[2];
    const uint32_t item = blockIdx.x;
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
        if (amp >= thr) {
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
void launch_init(hipStream_t st, uint64_t base, const uint32_t *ks, const uint64_t *seeds, uint32_t n, SideParams *out) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_init, dim3((2 * n + 63) / 64), dim3(64), 0, st, base, ks, seeds, n, out);
}
void launch_amp(hipStream_t st, int mode, const SideParams *sides, const uint32_t *sel, uint32_t n, const GateKernel &K,
                float thr, const uint32_t *kmap, uint32_t *out_idx, float *out_amp, uint32_t *count, uint32_t cap,
                float *all_amp) {
    if (n == 0) return;
    if (mode == 0)
        hipLaunchKernelGGL(k_amp<0>, dim3(n), dim3(128), 0, st, sides, sel, n, K, thr, kmap, out_idx, out_amp, count, cap,
                           all_amp);
    else if (mode == 1)
        hipLaunchKernelGGL(k_amp<1>, dim3(n), dim3(128), 0, st, sides, sel, n, K, thr, kmap, out_idx, out_amp, count, cap,
                           all_amp);
    else
        hipLaunchKernelGGL(k_amp<2>, dim3(n), dim3(128), 0, st, sides, sel, n, K, thr, kmap, out_idx, out_amp, count, cap,
                           all_amp);
}

}  // namespace dsx
