// CPU ORACLE - test infrastructure only. C entry points for the Python tests (ctypes) and for the
// bench's cpu_baseline leg. Nothing in the product links or loads this library.
#include "oracle.hpp"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace oracle;

namespace {
// Unnormalised gate blur weights expf(-0.5 d^2 / sig^2) in float, as gate.cuh:216 forms them, but
// evaluated with the host libm (the GPU library receives host-computed weights the same way).
void gate_kern(int R, float sig, float *k) {
    for (int t = 0; t <= 2 * R; t++) {
        const float d = float(t - R);
        k[t] = expf(-0.5f * d * d / (sig * sig));
    }
}
struct FloodOut { uint64_t cells, sampled; int32_t minx, maxx, minz, maxz, edge, border; };
}  // namespace

extern "C" {

uint64_t or_mix(uint64_t x) { return mix(x); }

double or_phase(uint64_t seed, double *fy) { return phase_t(seed, fy); }

void or_gate_kern(int R, float sig, float *k) { gate_kern(R, sig, k); }

void or_init_side(uint64_t seed, int side, uint8_t *perm, double *o6, int *iy) {
    PerlinSide p;
    init_side(seed, side, p);
    std::memcpy(perm, p.perm, 256);
    o6[0] = p.ox; o6[1] = p.oy; o6[2] = p.oz; o6[3] = p.fy; o6[4] = p.vy; o6[5] = 0;
    *iy = p.iy;
}

// Amplitude A + B of the GPU gate's stages: (G, R, box, sig) = (256, 4, 1, 1) pre-gate,
// (512, 7, 0, 2) full gate, (256, 4, 0, 1) legacy pre-gate.
float or_gate_amp(uint64_t seed, int G, int R, int box, float sig) {
    float k[32];
    gate_kern(R, sig, k);
    return gate_amp(seed, G, R, box != 0, k);
}
float or_gate_side(uint64_t seed, int side, int G, int R, int box, float sig) {
    float k[32];
    gate_kern(R, sig, k);
    PerlinSide p;
    init_side(seed, side, p);
    return gate_side_max(p, G, R, box != 0, k);
}

// Value grid (G^2 floats) of one side of the GPU gate, before blurring.
void or_gate_grid(uint64_t seed, int side, int G, int box, float *V) {
    PerlinSide p;
    init_side(seed, side, p);
    gate_grid(p, G, box != 0, V);
}
float or_gate_blur_max(const float *V, int G, int R, const float *kern) { return gate_blur_max(V, G, R, kern); }
// Side `side` of the temperature octave 0 tabulated as lowfield.hpp tabulate_octave does, with octave
// weight w (for the identity check against the full gate's grid).
void or_tabulate_side(uint64_t seed, int side, double w, float *out) {
    PerlinSide ps;
    init_side(seed, side, ps);
    Perlin p;
    std::memcpy(p.perm, ps.perm, 256);
    p.ox = ps.ox; p.oy = ps.oy; p.oz = ps.oz; p.fy = ps.fy; p.vy = ps.vy; p.iy = ps.iy;
    p.frequency = 1; p.weight = 1;
    std::vector<float> t;
    tabulate(p, w, t);
    std::memcpy(out, t.data(), t.size() * 4);
}
// Direct improved-noise sample of side `side` of temperature octave 0 at lattice coords (u, 0, v).
double or_perlin_side(uint64_t seed, int side, double u, double v) {
    PerlinSide ps;
    init_side(seed, side, ps);
    Perlin p;
    std::memcpy(p.perm, ps.perm, 256);
    p.ox = ps.ox; p.oy = ps.oy; p.oz = ps.oz; p.fy = ps.fy; p.vy = ps.vy; p.iy = ps.iy;
    p.frequency = 1; p.weight = 1;
    return p.sample(u, 0, v);
}

// GPU-gate semantics over seed indices [start, start+count): phase <= (double)maxph, then the
// 4-parity pre-gate >= pthr, then the full gate >= amp. Survivors sorted by seed index.
// counts = {phase survivors, pre-gate survivors, full survivors}. Returns the number written.
int64_t or_gate_run(uint64_t start, uint64_t count, float maxph, float pthr, float amp, int threads,
                    uint64_t *out_idx, float *out_amp, int64_t cap, uint64_t *counts) {
    float kp[32], kf[32];
    gate_kern(4, 1.0f, kp);
    gate_kern(7, 2.0f, kf);
    std::atomic<uint64_t> next(0), n1(0), n2(0);
    std::mutex mu;
    std::vector<std::pair<uint64_t, float>> pass;
    const uint64_t chunk = 4096;
    auto work = [&] {
        for (;;) {
            const uint64_t b = next.fetch_add(chunk);
            if (b >= count) break;
            const uint64_t e = std::min(count, b + chunk);
            for (uint64_t k = b; k < e; k++) {
                const uint64_t seed = mix(start + k);
                double fy[2];
                if (!(phase_t(seed, fy) <= double(maxph))) continue;
                n1++;
                if (!(gate_amp(seed, 256, 4, true, kp) >= pthr)) continue;
                n2++;
                const float a = gate_amp(seed, 512, 7, false, kf);
                if (a >= amp) {
                    std::lock_guard<std::mutex> g(mu);
                    pass.push_back({k, a});
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); t++) pool.emplace_back(work);
    for (auto &t : pool) t.join();
    std::sort(pass.begin(), pass.end());
    counts[0] = n1; counts[1] = n2; counts[2] = pass.size();
    int64_t n = std::min<int64_t>(cap, int64_t(pass.size()));
    for (int64_t i = 0; i < n; i++) { out_idx[i] = pass[i].first; if (out_amp) out_amp[i] = pass[i].second; }
    return n;
}

// Phase survivors only (k_phase semantics): indices k with phase(mix(start+k)) <= (double)maxph.
int64_t or_phase_list(uint64_t start, uint64_t count, float maxph, uint64_t *out_idx, int64_t cap) {
    int64_t n = 0;
    for (uint64_t k = 0; k < count; k++) {
        double fy[2];
        if (phase_t(mix(start + k), fy) <= double(maxph)) {
            if (n < cap) out_idx[n] = k;
            n++;
        }
    }
    return n;
}

// CPU-default pre-gate estimate (lowfield.hpp amplitude_sub on both sides).
float or_amplitude_sub(uint64_t seed) {
    Climate S(seed);
    const Noise &T = S.noise[Temperature];
    return amplitude_sub(T.oct[0][0], T.factor * T.oct[0][0].weight) + amplitude_sub(T.oct[1][0], T.factor * T.oct[1][0].weight);
}

// Table inputs (lowfield.hpp tab_setup) of table k in {TA, TB, CA, CB, EA, EB}: perm 256 B,
// col 512 B, T 512 floats, i0[2], k0[2].
void or_tab_setup(uint64_t seed, int k, uint8_t *perm, uint8_t *col, float *T, int *i0k0) {
    Climate S(seed);
    const Noise *nz[3] = {&S.noise[Temperature], &S.noise[Continentalness], &S.noise[Erosion]};
    const Noise &nn = *nz[k / 2];
    TabIn in;
    tab_setup(nn.oct[k & 1][0], nn.factor * nn.oct[k & 1][0].weight, in);
    std::memcpy(perm, in.perm, 256);
    std::memcpy(col, in.col, 512);
    std::memcpy(T, in.T, 512 * 4);
    i0k0[0] = in.i0[0]; i0k0[1] = in.i0[1]; i0k0[2] = in.k0[0]; i0k0[3] = in.k0[1];
}

// The 6 blurred period tables of a seed, 512*512 floats each, in out[6*512*512]. blurred=0 skips the blur.
void or_tables(uint64_t seed, double sigma, double sigmac, int blurred, float *out) {
    Climate S(seed);
    const Noise *nz[3] = {&S.noise[Temperature], &S.noise[Continentalness], &S.noise[Erosion]};
    Blur bt(sigma), bc(sigmac);
    std::vector<float> t, tmp;
    for (int k = 0; k < 6; k++) {
        const Noise &nn = *nz[k / 2];
        tabulate(nn.oct[k & 1][0], nn.factor * nn.oct[k & 1][0].weight, t);
        if (blurred) (k < 2 ? bt : bc).apply(t, tmp);
        std::memcpy(out + size_t(k) * TN * TN, t.data(), size_t(TN) * TN * 4);
    }
}
void or_blur_kernel(double sigma, float *k, int *r) {
    Blur b(sigma);
    *r = b.r;
    std::memcpy(k, b.k.data(), b.k.size() * 4);
}
void or_transmap(int which, int16_t *ib, int8_t *lo, int8_t *hi) {
    const TransMap &m = which ? tmap_c() : tmap_t();
    std::memcpy(ib, m.ib, sizeof m.ib);
    std::memcpy(lo, m.lo, sizeof m.lo);
    std::memcpy(hi, m.hi, sizeof m.hi);
}

// Raw scan candidates of a seed in (j, i, tz, tx) order: 6 floats/ints per row {s, t, c, e, x, z}
// (x, z stored as int bits). maxtab gets {maxta, maxtb}. Returns the total count.
int64_t or_scan(uint64_t seed, double ta, double tt, double wc, double we, void *out, int64_t cap, float *maxtab) {
    Params P;
    P.ta = ta; P.t = tt; P.wc = wc; P.we = we;
    std::vector<float> tabs[6];
    std::vector<float> all(size_t(6) * TN * TN);
    or_tables(seed, P.sigma, P.sigmac, 1, all.data());
    for (int k = 0; k < 6; k++) tabs[k].assign(all.begin() + size_t(k) * TN * TN, all.begin() + size_t(k + 1) * TN * TN);
    maxtab[0] = *std::max_element(tabs[0].begin(), tabs[0].end());
    maxtab[1] = *std::max_element(tabs[1].begin(), tabs[1].end());
    std::vector<Raw> raw;
    scan_raw(tabs, P, raw);
    const int64_t n = std::min<int64_t>(cap, int64_t(raw.size()));
    std::memcpy(out, raw.data(), size_t(n) * sizeof(Raw));
    return int64_t(raw.size());
}

double or_noise(uint64_t seed, int channel, double x, double y, double z) {
    Climate S(seed);
    return S.noise[channel].sample(x, y, z);
}
// Batched noise samples at integer quarts (y = 0): out[i] = N_channel(xs[i], 0, zs[i]).
void or_noise_batch(uint64_t seed, int channel, const int32_t *xs, const int32_t *zs, int64_t n, double *out) {
    Climate S(seed);
    for (int64_t i = 0; i < n; i++) out[i] = S.noise[channel].sample(xs[i], 0, zs[i]);
}
void or_climate(uint64_t seed, int x, int z, int64_t *q) {
    Climate S(seed);
    S.climate(x, 64, z, q);
}
int or_biome(uint64_t seed, int x, int z) { Climate S(seed); return S.biome(x, z); }
int or_classify(const int64_t *q) { return classify(q); }

// The biome map of tools/biomemap.cpp:11-45 into out[n * n] (n = 2 * half / step + 1): biome id at quart
// ((X - half + i * step) >> 2, (Z - half + j * step) >> 2), row j major; 255 on the 4-connected component
// of the centre over desert_state != 0 where desert_state == 2. Returns n, or -1 if the centre is not
// desert (out then holds the plain biome map).
int or_biomemap(uint64_t seed, int X, int Z, int half, int step, uint8_t *out) {
    const int n = 2 * half / step + 1;
    Climate S(seed);
    std::vector<uint8_t> st(size_t(n) * n), seen(size_t(n) * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const int qx = (X - half + i * step) >> 2, qz = (Z - half + j * step) >> 2;
            out[size_t(j) * n + i] = uint8_t(S.biome(qx, qz));
            st[size_t(j) * n + i] = uint8_t(S.desert_state(qx, qz));
        }
    const size_t c0 = size_t(n / 2) * n + n / 2;
    if (st[c0] != 2) return -1;
    // breadth-first here (the tool walks depth-first): the marked set is the component either way
    std::vector<size_t> q{c0};
    seen[c0] = 1;
    for (size_t h = 0; h < q.size(); h++) {
        const size_t c = q[h];
        if (st[c] == 2) out[c] = 255;
        const int ci = int(c % n), cj = int(c / n);
        for (int k = 0; k < 4; k++) {
            const int i = ci + (k == 0) - (k == 1), j = cj + (k == 2) - (k == 3);
            if (i < 0 || j < 0 || i >= n || j >= n) continue;
            const size_t d = size_t(j) * n + i;
            if (!seen[d] && st[d]) {
                seen[d] = 1;
                q.push_back(d);
            }
        }
    }
    return n;
}
int or_grid_lookup(const int64_t *q) { return grid().lookup(q); }
float or_spline(const float *v) { return spline(0, v); }

// Predicate kinds: 0 desert_state (bridged), 1 hot (T >= 0.55), 2 proxy(pcut), 3 desert() as 2/0.
void or_pred_batch(uint64_t seed, int kind, float pcut, const int32_t *xs, const int32_t *zs, int64_t n, int8_t *out) {
    Climate S(seed);
    for (int64_t i = 0; i < n; i++) {
        const int x = xs[i], z = zs[i];
        out[i] = int8_t(kind == 0 ? S.desert_state(x, z) : kind == 1 ? S.hot(x, z) : kind == 2 ? S.proxy(x, z, pcut)
                                                                                  : (S.desert(x, z) ? 2 : 0));
    }
}

// Bounded bridged flood with predicate `kind` (see or_pred_batch). If state_out is non-null it gets
// the (2*half+1)^2 final window state.
void or_flood(uint64_t seed, int kind, float pcut, int qx, int qz, int step, int half, int conn8, FloodOut *out,
              uint8_t *state_out) {
    Climate S(seed);
    std::vector<uint8_t> state;
    std::vector<int> q;
    Area a;
    if (kind == 0) a = flood([&](int x, int z) { return S.desert_state(x, z); }, qx, qz, step, half, conn8, state, q);
    else if (kind == 1) a = flood([&](int x, int z) { return S.hot(x, z); }, qx, qz, step, half, conn8, state, q);
    else if (kind == 2) a = flood([&](int x, int z) { return S.proxy(x, z, pcut); }, qx, qz, step, half, conn8, state, q);
    else a = flood([&](int x, int z) { return S.desert(x, z) ? 2 : 0; }, qx, qz, step, half, conn8, state, q);
    out->cells = a.cells; out->sampled = a.sampled;
    out->minx = a.minx; out->maxx = a.maxx; out->minz = a.minz; out->maxz = a.maxz;
    out->edge = a.edge; out->border = a.border;
    if (state_out) std::memcpy(state_out, state.data(), state.size());
}

// Whole pipeline over seed indices [start, start+count) (gated=0, seeds = mix(start+k)) or over an
// explicit seed list (gated=1). Writes PREFIX.csv, PREFIX.seeds.csv and PREFIX.json in the reference's
// formats (rows in completion order) and the CANDIDATE lines to PREFIX.cand. counters = {done, passed,
// scanned, sites}; seconds = wall time. params (optional): {max_phase, amp, amp_margin} and, with
// nparams >= 6, {exact_min, refine_mb, verify_mb} (the `exact_min`, `refine`, `verify` keys).
int or_run(uint64_t start, uint64_t count, const uint64_t *seeds, int64_t nseeds, int gated, int threads,
           const char *prefix, uint64_t *counters, double *seconds, const double *params, int nparams) {
    Params P;
    if (params && nparams >= 3) { P.max_phase = params[0]; P.amp = params[1]; P.amp_margin = params[2]; }
    if (params && nparams >= 6) { P.exact_min = params[3]; P.refine_mb = params[4]; P.verify_mb = params[5]; }
    std::string pre(prefix);
    FILE *out = std::fopen((pre + ".csv").c_str(), "w");
    FILE *sout = std::fopen((pre + ".seeds.csv").c_str(), "w");
    FILE *cout = std::fopen((pre + ".cand").c_str(), "w");
    if (!out || !sout || !cout) return 2;
    std::fprintf(out, "seed,fyA,fyB,maxta,maxtb,rank,x,z,score,t0,c0,e0,proxy,s8,s4,edge,hot\n");
    std::fprintf(sout, "seed,fyA,fyB,maxta,maxtb,candidates,sites,best\n");
    std::atomic<uint64_t> next(0), done(0), passed(0), scanned(0), nsites(0);
    std::mutex mu;
    const uint64_t total = gated ? uint64_t(nseeds) : count;
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&] {
        for (;;) {
            const uint64_t k = next++;
            if (k >= total) break;
            const uint64_t seed = gated ? seeds[k] : mix(start + k);
            SeedResult r = run_seed(seed, P, gated != 0);
            done++;
            if (!r.phase_pass) continue;
            passed++;
            if (!r.pre_pass) continue;
            if (r.scanned) { scanned++; nsites += r.nsites; }
            const std::string rows = r.scanned ? csv_rows(r, P) : std::string();
            const std::string srow = seeds_row(r);
            std::lock_guard<std::mutex> g(mu);
            std::fputs(rows.c_str(), out);
            std::fputs(srow.c_str(), sout);
            for (auto &c : r.candidates) std::fprintf(cout, "%s\n", c.c_str());
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); t++) pool.emplace_back(work);
    for (auto &t : pool) t.join();
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fclose(out);
    std::fclose(sout);
    std::fclose(cout);
    counters[0] = done; counters[1] = passed; counters[2] = scanned; counters[3] = nsites;
    *seconds = sec;
    return 0;
}

}  // extern "C"
