// Per-seed climate initialisation and batched evaluation kernels (noise, climate vector, predicates).
#include "climate.h"
#include "climate_api.h"

namespace dsx {

// One lane per octave slot (46 per seed; seed s goes to out[dst[s]]): NormalNoise::init (standalone.hpp:123-139) for the slot's
// channel and side, then Perlin::init (standalone.hpp:55-61) with the permutation shuffled in a
// lane-private LDS slot.
constexpr int CSTRIDE = 260;
__global__ __launch_bounds__(64) void k_climinit(const uint64_t *seeds, const uint32_t *dst, uint32_t n, ClimSeed *out) {
    __shared__ uint32_t slot32[64 * CSTRIDE / 4];
    uint8_t *slots = reinterpret_cast<uint8_t *>(slot32);
    const int lane = threadIdx.x;
    const uint32_t g = blockIdx.x * 64u + lane;
    const bool active = g < CLIM_NOCT * n;
    uint8_t *my = slots + lane * CSTRIDE;
    double ox = 0, oy = 0, oz = 0, freq = 0, weight = 0;
    // seeds and dst may be in pinned host memory: each is read once per lane, dst passed on by shuffle
    const uint32_t my_dst = active ? dst[g / CLIM_NOCT] : 0u;
    if (active) {
        const uint32_t s = g / CLIM_NOCT, slot = g % CLIM_NOCT;
        int ch = 5;
        while (clim_base(ch) > int(slot)) ch--;
        const int n_ch = clim_n(ch), rel = int(slot) - clim_base(ch), side = rel / n_ch, j = rel % n_ch;
        const dsx_noise_cfg &cfg = DSX_NOISE[ch];
        Xoro w = Xoro::seeded(seeds[s]);
        const uint64_t a = w.next(), b = w.next();
        Xoro r = Xoro::raw(a ^ cfg.fork_a, b ^ cfg.fork_b);
        uint64_t ra = r.next(), rb = r.next();
        if (side) { ra = r.next(); rb = r.next(); }
        double f = ldexp(1., cfg.first);
        double wt = double(1 << (cfg.len - 1)) / ((1 << cfg.len) - 1);
        int i = 0, cnt = -1;
        for (; i < cfg.len; i++, f *= 2, wt *= .5)
            if (cfg.amp[i] != 0 && ++cnt == j) break;
        const uint64_t *salt = DSX_OCT_SALT[cfg.first + i + 10];
        Xoro q = Xoro::raw(ra ^ salt[0], rb ^ salt[1]);
        ox = q.uniform() * 256;
        oy = q.uniform() * 256;
        oz = q.uniform() * 256;
        freq = f;
        weight = wt * cfg.amp[i];
        uint32_t *my32 = reinterpret_cast<uint32_t *>(my);
        for (int k = 0; k < 64; k++) my32[k] = 0x03020100u + 0x04040404u * uint32_t(k);
        for (int k = 0; k < 256; k++) {
            const int m = k + int(q.bounded(uint32_t(256 - k)));
            const uint8_t t = my[k];
            my[k] = my[m];
            my[m] = t;
        }
    }
    __syncthreads();
    const uint32_t g0 = blockIdx.x * 64u;
    for (int k = 0; k < 64; k++) {
        const uint32_t gk = g0 + k;
        if (gk >= CLIM_NOCT * n) break;
        const uint32_t dk = __shfl(my_dst, k);
        reinterpret_cast<uint32_t *>(out[dk].oct[gk % CLIM_NOCT].perm)[lane] = slot32[k * (CSTRIDE / 4) + lane];
    }
    if (active) {
        ClimSeed &cs = out[my_dst];
        ClimOct &o = cs.oct[g % CLIM_NOCT];
        o.ox = ox;
        o.oy = oy;
        o.oz = oz;
        o.freq = freq;
        o.weight = weight;
        o.iy = int(floor(oy));
        o.fy = oy - floor(oy);
        o.vy = fade(o.fy);
        if (g % CLIM_NOCT == 0) {
            GradTab &G = cs.g;
            for (int h = 0; h < 16; h++) {
                G.xy[h] = make_double2(grad_component(0, h), grad_component(1, h));
                G.z[h] = grad_component(2, h);
            }
            for (int ch = 0; ch < 6; ch++) {
                const dsx_noise_cfg &c = DSX_NOISE[ch];
                int lo = 0, hi = c.len - 1;
                while (c.amp[lo] == 0) lo++;
                while (c.amp[hi] == 0) hi--;
                const int span = hi - lo + 1;
                cs.factor[ch] = (5. * span) / (3. * (span + 1));
            }
        }
    }
}

// Copy one seed's climate into LDS (all threads of the block participate).
__device__ __forceinline__ void load_climate(ClimSeed &dst, const ClimSeed &src) {
    const uint4 *s = reinterpret_cast<const uint4 *>(&src);
    uint4 *d = reinterpret_cast<uint4 *>(&dst);
    for (int k = threadIdx.x; k < int(sizeof(ClimSeed) / 16); k += blockDim.x) d[k] = s[k];
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_noise(const ClimSeed *cs, int ch, const double *pts, uint32_t n, double *out) {
    __shared__ ClimSeed C;
    load_climate(C, *cs);
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < n) out[k] = noise_sample(C, ch, pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
}

__global__ __launch_bounds__(256) void k_noise32(const ClimSeed *cs, int ch, const double *pts, uint32_t n, double *out) {
    __shared__ ClimSeed C;
    load_climate(C, *cs);
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < n) out[k] = noise_sample32(C, ch, pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
}

__global__ __launch_bounds__(256) void k_climate(const ClimSeed *cs, const int32_t *xs, const int32_t *zs, uint32_t n,
                                                 int64_t *q) {
    __shared__ ClimSeed C;
    load_climate(C, *cs);
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < n) climate_q(C, xs[k], 64, zs[k], q + 6 * size_t(k));
}

__global__ __launch_bounds__(256) void k_pred(const ClimSeed *cs, SurfGrid g, int kind, float cut, const int32_t *xs,
                                              const int32_t *zs, uint32_t n, int8_t *out, int pre) {
    __shared__ ClimSeed C;
    load_climate(C, *cs);
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < n) out[k] = int8_t(predicate_mode(pre, kind, C, g, xs[k], zs[k], cut));
}

// Points of many seeds: the climate stays in global memory (L2), one thread per point.
__global__ __launch_bounds__(256) void k_points(const ClimSeed *cs, SurfGrid g, const uint32_t *sidx, const int32_t *xs,
                                                const int32_t *zs, uint32_t n, int what, float cut, int8_t *out8,
                                                double *outd, int pre) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const ClimSeed &C = cs[sidx[k]];
    if (what == 16) outd[k] = noise_sample(C, DSX_CH_T, xs[k], 0, zs[k]);
    else if (what == 17) outd[k] = proxy_prob(C, xs[k], zs[k]);
    else if (what == 20) outd[k] = noise_octaves(C, DSX_CH_C, xs[k], zs[k], 3);   // desert_proxy.hpp:14-24
    else if (what == 21) outd[k] = noise_octaves(C, DSX_CH_E, xs[k], zs[k], 2);
    else if (what == 19) {   // surface.hpp:79-83
        int64_t q[6];
        climate_q(C, xs[k], 64, zs[k], q);
        const int b = grid_lookup(g, q);
        out8[k] = int8_t(b >= 0 ? b : classify(q));
    } else out8[k] = int8_t(predicate_mode(pre, what, C, g, xs[k], zs[k], cut));
}

// Hottest of the 7 x 7 points around each site (one wave per site, lane l < 49 = (dz + 3) * 7 + dx + 3):
// T from global memory (L2) as k_points does; the wave keeps the largest T, the lowest index on ties,
// which is the first maximum of the reference's dz-then-dx scan with a strict greater-than.
__global__ __launch_bounds__(256) void k_hottest(const ClimSeed *cs, const uint32_t *sidx, const int32_t *xs,
                                                 const int32_t *zs, uint32_t n, int8_t *out8, double *outd) {
    const uint32_t site = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (site >= n) return;
    const ClimSeed &C = cs[sidx[site]];
    double t = -1e300;
    int idx = lane;
    if (lane < 49) t = noise_sample(C, DSX_CH_T, xs[site] + (lane % 7 - 3) * 256, 0, zs[site] + (lane / 7 - 3) * 256);
    for (int o = 32; o > 0; o >>= 1) {
        const double ot = __shfl_xor(t, o);
        const int oi = __shfl_xor(idx, o);
        if (ot > t || (ot == t && oi < idx)) {
            t = ot;
            idx = oi;
        }
    }
    if (lane == 0) {
        outd[site] = t;
        out8[site] = int8_t(idx);
    }
}

// ------------------------------------------------------------------ proxy-table Monte Carlo (tools/dfunc.cpp)
// Octaves [lo, hi) of channel ch in one running sum over both sides, times the factor (tools/dfunc.cpp:12-21).
__device__ __forceinline__ double octave_range(const ClimSeed &C, int ch, double x, double z, int lo, int hi) {
    const int n = clim_n(ch);
    double s = 0;
    for (int side = 0; side < 2; side++) {
        const double sc = side ? 337.0 / 331.0 : 1.0;
#pragma unroll 1
        for (int k = lo; k < hi && k < n; k++) {
            const ClimOct &o = C.oct[oct_slot<ClimSeed>(ch, side, k)];
            s += o.weight * perlin_sample(C.g, o, x * sc * o.freq, 0, z * sc * o.freq);
        }
    }
    return s * C.factor[ch];
}

// tools/dfunc.cpp:52-55: a try is a hot point when T at the shifted position is >= 0.55.
__global__ __launch_bounds__(256) void k_dfunc_hot(const ClimSeed *cs, const uint32_t *sidx, const int32_t *xs,
                                                   const int32_t *zs, uint32_t n, uint8_t *hot) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const ClimSeed &C = cs[sidx[k]];
    double px, pz;
    shifted(C, xs[k], zs[k], px, pz);
    hot[k] = noise_sample(C, DSX_CH_T, px, 0, pz) >= 0.55;
}

// tools/dfunc.cpp:56-65 for one hot point: surface biome, MC-Measure bridge to desert, low-octave C/E bins of
// both sets, counted into hist[((set * 40 + ci) * 40 + ei) * 3 + {total, desert, passable}] (integer atomics,
// so the histogram does not depend on the order of the points).
__device__ __forceinline__ int dfunc_bin(double v, double lo) {
    const int i = int(floor((v + lo) / 0.05));
    return i < 0 ? 0 : (i > 39 ? 39 : i);
}
__global__ __launch_bounds__(256) void k_dfunc_rec(const ClimSeed *cs, SurfGrid g, const uint32_t *sidx,
                                                   const int32_t *xs, const int32_t *zs, uint32_t n,
                                                   unsigned long long *hist) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const ClimSeed &C = cs[sidx[k]];
    const int x = xs[k], z = zs[k];
    double px, pz;
    shifted(C, x, z, px, pz);
    int64_t q[6];
    climate_q(C, x, 64, z, q);
    int b = grid_lookup(g, q);
    if (b < 0) b = classify(q);
    const bool des = b == 2;
    bool pass = des;
    if (!des && (b == 7 || b == 11) && q[5] <= 500 && q[5] >= -500) {   // surface.hpp:126-134
        q[5] = q[5] >= 0 ? 501 : -501;
        int u = grid_lookup(g, q);
        if (u < 0) u = classify(q);
        pass = u == 2;
    }
    const double c0 = octave_range(C, DSX_CH_C, px, pz, 0, 1), c02 = c0 + octave_range(C, DSX_CH_C, px, pz, 1, 3);
    const double e0 = octave_range(C, DSX_CH_E, px, pz, 0, 1), e01 = e0 + octave_range(C, DSX_CH_E, px, pz, 1, 2);
    for (int set = 0; set < 2; set++) {
        const int i = dfunc_bin(set ? c02 : c0, 0.8), j = dfunc_bin(set ? e01 : e0, 1.0);
        unsigned long long *cell = hist + ((set * 40 + i) * 40 + j) * 3;
        atomicAdd(cell, 1ull);
        if (des) atomicAdd(cell + 1, 1ull);
        if (pass) atomicAdd(cell + 2, 1ull);
    }
}

void launch_dfunc_hot(hipStream_t st, const ClimSeed *cs, const uint32_t *sidx, const int32_t *xs, const int32_t *zs,
                      uint32_t n, uint8_t *hot) {
    if (n) hipLaunchKernelGGL(k_dfunc_hot, dim3((n + 255) / 256), dim3(256), 0, st, cs, sidx, xs, zs, n, hot);
}
void launch_dfunc_rec(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, const uint32_t *sidx, const int32_t *xs,
                      const int32_t *zs, uint32_t n, unsigned long long *hist) {
    if (n) hipLaunchKernelGGL(k_dfunc_rec, dim3((n + 255) / 256), dim3(256), 0, st, cs, g, sidx, xs, zs, n, hist);
}

void launch_climinit(hipStream_t st, const uint64_t *seeds, const uint32_t *slots, uint32_t n, ClimSeed *out) {
    if (n) hipLaunchKernelGGL(k_climinit, dim3((CLIM_NOCT * n + 63) / 64), dim3(64), 0, st, seeds, slots, n, out);
}
void launch_noise(hipStream_t st, const ClimSeed *cs, int ch, const double *pts, uint32_t n, double *out, int fp32) {
    if (!n) return;
    if (fp32) hipLaunchKernelGGL(k_noise32, dim3((n + 255) / 256), dim3(256), 0, st, cs, ch, pts, n, out);
    else hipLaunchKernelGGL(k_noise, dim3((n + 255) / 256), dim3(256), 0, st, cs, ch, pts, n, out);
}
void launch_climate(hipStream_t st, const ClimSeed *cs, const int32_t *xs, const int32_t *zs, uint32_t n, int64_t *q) {
    if (n) hipLaunchKernelGGL(k_climate, dim3((n + 255) / 256), dim3(256), 0, st, cs, xs, zs, n, q);
}
void launch_points(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, const uint32_t *sidx, const int32_t *xs,
                   const int32_t *zs, uint32_t n, int what, float cut, int8_t *out8, double *outd, int pre) {
    if (!n) return;
    if (what == 18) hipLaunchKernelGGL(k_hottest, dim3((n + 3) / 4), dim3(256), 0, st, cs, sidx, xs, zs, n, out8, outd);
    else hipLaunchKernelGGL(k_points, dim3((n + 255) / 256), dim3(256), 0, st, cs, g, sidx, xs, zs, n, what, cut, out8, outd,
                            pre);
}
void launch_pred(hipStream_t st, const ClimSeed *cs, const SurfGrid &g, int kind, float cut, const int32_t *xs,
                 const int32_t *zs, uint32_t n, int8_t *out, int pre) {
    if (n) hipLaunchKernelGGL(k_pred, dim3((n + 255) / 256), dim3(256), 0, st, cs, g, kind, cut, xs, zs, n, out, pre);
}

}  // namespace dsx
