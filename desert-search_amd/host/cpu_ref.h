// The reference's own CPU arithmetic for the validation switches, header-only host code:
//   gscancheck=N   period tables + blur + translation scan on the CPU for the first N seeds, compared
//                  with the GPU scan (dsearch_desert.cpp:228-242, 289-296)
//   gpucheck=N     the FP64 DesertProxy on the CPU at every 7th cell of the first N hot zones' step-32
//                  grids, compared with the GPU predicate (dsearch_desert.cpp:367-371, 445-450)
//   tgate check=1  the full 512^2 amplitude in the CPU arithmetic of tgate.cu:14-39 beside the GPU's
// The pipeline never calls this; the switches print both sides, as the reference's do. Semantics follow
// standalone.hpp:55-82 and 123-151 (Perlin, NormalNoise), lowfield.hpp:14-55 and 105-139 (tab_setup,
// tabulate_octave, Blur), desert_proxy.hpp:14-32 (DesertProxy), dsearch_desert.cpp:98-113
// (TranslationMap) and 256-280 (scan, buckets), with the same operations in the same order (host code,
// built without FMA contraction).
#pragma once
#include "../csrc/dsx_rng.h"
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <utility>
#include <vector>

namespace dsx::cpu {

// ------------------------------------------------------------------ noise
struct Octave {
    uint8_t perm[256];
    double ox, oy, oz, freq, weight, fy, vy;
    int iy;
};

inline void octave_init(Xoro r, double f, double w, Octave &o) {
    o.ox = r.uniform() * 256;
    o.oy = r.uniform() * 256;
    o.oz = r.uniform() * 256;
    o.freq = f;
    o.weight = w;
    for (int i = 0; i < 256; i++) o.perm[i] = uint8_t(i);
    for (int i = 0; i < 256; i++) std::swap(o.perm[i], o.perm[i + int(r.bounded(uint32_t(256 - i)))]);
    o.iy = int(std::floor(o.oy));
    o.fy = o.oy - std::floor(o.oy);
    o.vy = fade(o.fy);
}

inline double lerp(double t, double a, double b) { return a + t * (b - a); }

inline double perlin(const Octave &o, double x, double y, double z) {
    x += o.ox;
    z += o.oz;
    const double bx = std::floor(x), bz = std::floor(z);
    const int ix = int(bx), iz = int(bz);
    int jy = o.iy;
    x -= bx;
    z -= bz;
    double dy = o.fy, ty = o.vy;
    if (y != 0) {
        y += o.oy;
        const double by = std::floor(y);
        jy = int(by);
        dy = y - by;
        ty = fade(dy);
    }
    const uint8_t *P = o.perm;
    auto h = [&](int dx, int dyy, int dz) { return unsigned(P[(P[(P[(ix + dx) & 255] + jy + dyy) & 255] + iz + dz) & 255]); };
    const double tx = fade(x), tz = fade(z);
    const double lo = lerp(ty, lerp(tx, grad(h(0, 0, 0), x, dy, z), grad(h(1, 0, 0), x - 1, dy, z)),
                           lerp(tx, grad(h(0, 1, 0), x, dy - 1, z), grad(h(1, 1, 0), x - 1, dy - 1, z)));
    const double hi = lerp(ty, lerp(tx, grad(h(0, 0, 1), x, dy, z - 1), grad(h(1, 0, 1), x - 1, dy, z - 1)),
                           lerp(tx, grad(h(0, 1, 1), x, dy - 1, z - 1), grad(h(1, 1, 1), x - 1, dy - 1, z - 1)));
    return lerp(tz, lo, hi);
}

struct Noise {
    Octave oct[2][9];
    int n = 0;
    double factor = 0;
};

// Channel ch of the world seed (Sampler: the world RNG's first two words, then the channel's fork).
inline void noise_init(uint64_t seed, int ch, Noise &N) {
    const dsx_noise_cfg &c = DSX_NOISE[ch];
    Xoro w = Xoro::seeded(seed);
    const uint64_t a = w.next(), b = w.next();
    Xoro r = Xoro::raw(a ^ c.fork_a, b ^ c.fork_b);
    for (int side = 0; side < 2; side++) {
        const uint64_t ra = r.next(), rb = r.next();
        N.n = 0;
        double f = std::ldexp(1., c.first);
        double wt = double(1 << (c.len - 1)) / ((1 << c.len) - 1);
        for (int i = 0; i < c.len; i++, f *= 2, wt *= .5)
            if (c.amp[i] != 0) {
                const uint64_t *salt = DSX_OCT_SALT[c.first + i + 10];
                octave_init(Xoro::raw(ra ^ salt[0], rb ^ salt[1]), f, wt * c.amp[i], N.oct[side][N.n++]);
            }
    }
    int lo = 0, hi = c.len - 1;
    while (c.amp[lo] == 0) lo++;
    while (c.amp[hi] == 0) hi--;
    const int span = hi - lo + 1;
    N.factor = (5. * span) / (3. * (span + 1));
}

inline double noise(const Noise &N, double x, double y, double z) {
    double sums[2] = {0, 0};
    for (int side = 0; side < 2; side++) {
        const double scale = side ? 337. / 331. : 1.;
        const double px = x * scale, py = y * scale, pz = z * scale;
        for (int j = 0; j < N.n; j++) {
            const Octave &o = N.oct[side][j];
            sums[side] += o.weight * perlin(o, px * o.freq, py * o.freq, pz * o.freq);
        }
    }
    return (sums[0] + sums[1]) * N.factor;
}

// DesertProxy::octaves: one running sum over both sides of octaves [lo, hi), times the factor.
inline double proxy_octaves(const Noise &N, double x, double z, int lo, int hi) {
    double s = 0;
    for (int side = 0; side < 2; side++) {
        const double sc = side ? 337.0 / 331.0 : 1.0;
        for (int k = lo; k < hi && k < N.n; k++) {
            const Octave &p = N.oct[side][k];
            s += p.weight * perlin(p, x * sc * p.freq, 0, z * sc * p.freq);
        }
    }
    return s * N.factor;
}

// DesertProxy::operator(): 2 if T >= 0.55 and the proxy table at the low-octave (C, E) bin >= cut, else 0.
template <class Table>
inline int proxy_pred(const Noise &T, const Noise &C, const Noise &E, const Table &prob, float cut, int x, int z) {
    float pr = 0;
    if (!(noise(T, x, 0, z) < 0.55)) {
        const double c = proxy_octaves(C, x, z, 0, 3), e = proxy_octaves(E, x, z, 0, 2);
        const int i = std::clamp(int(std::floor((c + 0.8) / 0.05)), 0, 39);
        const int j = std::clamp(int(std::floor((e + 1.0) / 0.05)), 0, 39);
        pr = prob[i][j];
    }
    return pr >= cut ? 2 : 0;
}

// ------------------------------------------------------------------ period tables, blur, scan
constexpr int TN = 512;

// t[j][i] = w * perlin(i / 2, 0, j / 2) of octave o as the sum of 8 float corner terms.
inline void tabulate(const Octave &o, double w, std::vector<float> &t) {
    const uint8_t *P = o.perm;
    uint8_t col[2][256];
    for (int x = 0; x < 256; x++)
        for (int dy = 0; dy < 2; dy++) col[dy][x] = P[(P[x] + o.iy + dy) & 255];
    int i0[2], k0[2];
    double fx[2], fz[2];
    for (int par = 0; par < 2; par++) {
        const double bx = par * 0.5 + o.ox, bz = par * 0.5 + o.oz;
        i0[par] = int(std::floor(bx));
        fx[par] = bx - std::floor(bx);
        k0[par] = int(std::floor(bz));
        fz[par] = bz - std::floor(bz);
    }
    float T[2][2][8][16];
    for (int px = 0; px < 2; px++)
        for (int pz = 0; pz < 2; pz++) {
            const double tx = fade(fx[px]), tz = fade(fz[pz]), ty = o.vy;
            for (int c = 0; c < 8; c++) {
                const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                const double wc = (dx ? tx : 1 - tx) * (dy ? ty : 1 - ty) * (dz ? tz : 1 - tz);
                for (unsigned h = 0; h < 16; h++)
                    T[px][pz][c][h] = float(w * wc * grad(h, fx[px] - dx, o.fy - dy, fz[pz] - dz));
            }
        }
    t.resize(size_t(TN) * TN);
    for (int j = 0; j < TN; j++) {
        const int pz = j & 1, iz = (j >> 1) + k0[pz];
        for (int i = 0; i < TN; i++) {
            const int px = i & 1, ix = (i >> 1) + i0[px];
            float acc = 0;
            for (int c = 0; c < 8; c++) {
                const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                acc += T[px][pz][c][P[(col[dy][(ix + dx) & 255] + iz + dz) & 255] & 15];
            }
            t[size_t(j) * TN + i] = acc;
        }
    }
}

// Periodic separable Gaussian blur: r = int(3 sigma + 1), weights float(exp(-i^2 / 2 sigma^2) / sum),
// each output a float sum over the taps in index order from 0 (rows, then columns).
struct Blur {
    std::vector<float> k;
    int r;
    explicit Blur(double sigma) {
        r = int(3 * sigma + 1);
        std::vector<double> kd(2 * r + 1);
        double s = 0;
        for (int i = -r; i <= r; i++) s += kd[i + r] = std::exp(-0.5 * i * i / (sigma * sigma));
        k.resize(2 * r + 1);
        for (int i = 0; i <= 2 * r; i++) k[i] = float(kd[i] / s);
    }
    void operator()(std::vector<float> &a) const {
        std::vector<float> h(size_t(TN) * TN);
        for (int j = 0; j < TN; j++)
            for (int i = 0; i < TN; i++) {
                float s = 0;
                for (int d = 0; d <= 2 * r; d++) s += k[d] * a[size_t(j) * TN + ((i + d - r) & (TN - 1))];
                h[size_t(j) * TN + i] = s;
            }
        for (int j = 0; j < TN; j++)
            for (int i = 0; i < TN; i++) {
                float s = 0;
                for (int d = 0; d <= 2 * r; d++) s += k[d] * h[size_t((j + d - r) & (TN - 1)) * TN + i];
                a[size_t(j) * TN + i] = s;
            }
    }
};

// A cell i of a table (i * 512 quarts within the 262144-quart period) translated by t periods -> the
// side-B cell of an octave whose side-A lattice unit is `lat` quarts; lo/hi: the translations inside the
// world (dsearch_desert.cpp:98-113).
struct TransMap {
    int16_t ib[TN][60];
    int8_t lo[TN], hi[TN];
    explicit TransMap(double lat) {
        for (int i = 0; i < TN; i++) {
            lo[i] = 60;
            hi[i] = -1;
            for (int t = 0; t < 60; t++) {
                const long x = long(i) * 512 + long(t - 30) * 262144;
                const bool ok = x >= -7499996 + 6000 && x <= 7499996 - 6000;
                double u = x * (337.0 / 331.0) / lat;
                u -= std::floor(u / 256.0) * 256.0;
                ib[i][t] = int16_t(int(u * (TN / 256.0) + 0.5) & (TN - 1));
                if (ok) {
                    lo[i] = int8_t(std::min<int>(lo[i], t));
                    hi[i] = int8_t(std::max<int>(hi[i], t));
                }
            }
        }
    }
};

struct Cand {
    float s, t, c, e;
    int x, z;
};

// The CPU tables and scan of one seed: maxima of the blurred side-A / side-B temperature tables and the
// best candidate per nms bucket in the order buckets first appear (dsearch_desert.cpp:246-280).
struct ScanResult {
    float maxta = 0, maxtb = 0;
    std::vector<Cand> cand;
};
inline uint64_t bucket_key(int x, int z, int nms) {
    return (uint64_t(uint32_t((x + 8000000) / nms)) << 32) | uint32_t((z + 8000000) / nms);
}
inline ScanResult scan_seed(uint64_t seed, double sigma, double sigmac, float ta, float tt, float wc, float we, int nms) {
    static const TransMap TMT(1024.0), TMC(512.0);
    Noise T, C, E;
    noise_init(seed, DSX_CH_T, T);
    noise_init(seed, DSX_CH_C, C);
    noise_init(seed, DSX_CH_E, E);
    const Blur blur(sigma), blurc(sigmac);
    std::vector<float> tab[6];
    const Noise *nz[3] = {&T, &C, &E};
    for (int k = 0; k < 6; k++) {
        const Noise &N = *nz[k / 2];
        const Octave &o = N.oct[k & 1][0];
        tabulate(o, N.factor * o.weight, tab[k]);
        (k < 2 ? blur : blurc)(tab[k]);
    }
    const std::vector<float> &TA = tab[0], &TB = tab[1], &CA = tab[2], &CB = tab[3], &EA = tab[4], &EB = tab[5];
    ScanResult R;
    R.maxta = *std::max_element(TA.begin(), TA.end());
    R.maxtb = *std::max_element(TB.begin(), TB.end());
    std::unordered_map<uint64_t, uint32_t> bucket;
    for (int j = 0; j < TN; j++)
        for (int i = 0; i < TN; i++) {
            const float sa = TA[size_t(j) * TN + i];
            if (sa < ta) continue;
            const int ca = ((2 * j) & (TN - 1)) * TN + ((2 * i) & (TN - 1));
            const float cA = CA[ca], eA = EA[ca];
            for (int tz = TMT.lo[j]; tz <= TMT.hi[j]; tz++) {
                const float *rT = &TB[size_t(TMT.ib[j][tz]) * TN];
                const float *rC = &CB[size_t(TMC.ib[j][tz]) * TN], *rE = &EB[size_t(TMC.ib[j][tz]) * TN];
                for (int tx = TMT.lo[i]; tx <= TMT.hi[i]; tx++) {
                    const float tv = sa + rT[TMT.ib[i][tx]];
                    if (tv < tt - 0.3f) continue;
                    const int cb = TMC.ib[i][tx];
                    const float cv = cA + rC[cb], ev = eA + rE[cb];
                    const float sc = tv + wc * cv + we * ev;
                    if (sc < tt) continue;
                    const int x = i * 512 + (tx - 30) * 262144, z = j * 512 + (tz - 30) * 262144;
                    auto [it, fresh] = bucket.try_emplace(bucket_key(x, z, nms), uint32_t(R.cand.size()));
                    if (fresh) R.cand.push_back({sc, tv, cv, ev, x, z});
                    else if (sc > R.cand[it->second].s) R.cand[it->second] = {sc, tv, cv, ev, x, z};
                }
            }
        }
    return R;
}

// ------------------------------------------------------------------ tgate check=1
// Full amplitude A + B of the temperature octave 0 in tgate.cu:14-39's CPU arithmetic: per side a G^2
// grid of float(w0 * perlin) evaluated in double at spacing 256/G, blurred with expf weights divided by
// their float sum, maximum; w0 = factor 1.25 * octave weight 32/63 * amplitude 1.5.
inline float gate_amp(uint64_t seed, int G, int R, float sig) {
    constexpr double W0 = 1.25 * (32.0 / 63.0) * 1.5;
    Noise Tn;
    noise_init(seed, DSX_CH_T, Tn);
    float amp = 0;
    for (int side = 0; side < 2; side++) {
        const Octave &p = Tn.oct[side][0];
        std::vector<float> t(size_t(G) * G), h(size_t(G) * G);
        for (int j = 0; j < G; j++)
            for (int i = 0; i < G; i++) {
                const double u = (G == 512 ? i * 0.5 : i) + p.ox, v = (G == 512 ? j * 0.5 : j) + p.oz;
                const int ix = int(std::floor(u)), iz = int(std::floor(v));
                const double fx = u - std::floor(u), fz = v - std::floor(v);
                const double tx = fade(fx), tz = fade(fz);
                double acc = 0;
                for (int c = 0; c < 8; c++) {
                    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                    const unsigned hsh = p.perm[(p.perm[(p.perm[(ix + dx) & 255] + p.iy + dy) & 255] + iz + dz) & 255];
                    const double wc = (dx ? tx : 1 - tx) * (dy ? p.vy : 1 - p.vy) * (dz ? tz : 1 - tz);
                    acc += wc * grad(hsh, fx - dx, p.fy - dy, fz - dz);
                }
                t[size_t(j) * G + i] = float(W0 * acc);
            }
        std::vector<float> k(2 * R + 1);
        float ks = 0;
        for (int d = -R; d <= R; d++) ks += k[d + R] = expf(-0.5f * d * d / (sig * sig));
        for (int j = 0; j < G; j++)
            for (int i = 0; i < G; i++) {
                float s = 0;
                for (int d = -R; d <= R; d++) s += k[d + R] * t[size_t(j) * G + ((i + d) & (G - 1))];
                h[size_t(j) * G + i] = s / ks;
            }
        float m = -9;
        for (int j = 0; j < G; j++)
            for (int i = 0; i < G; i++) {
                float s = 0;
                for (int d = -R; d <= R; d++) s += k[d + R] * h[size_t((j + d) & (G - 1)) * G + i];
                m = std::max(m, s / ks);
            }
        amp += m;
    }
    return amp;
}

}  // namespace dsx::cpu
