// CPU ORACLE - test infrastructure only.
//
// A clean-room CPU restatement of desert-search's per-seed pipeline, used ONLY by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker. The product (the HIP
// library under desert-search_amd/) never links, loads or calls anything here.
//
// Each function cites the reference file:line whose semantics it restates. The arithmetic keeps the
// reference's association order and float/double casts exactly, built with -ffp-contract=off, so
// results are bit-identical to the reference's CPU code (checked against the reference binaries'
// own outputs by tests/test_oracle_golden.py).
//
// Parity pinning (what each part is checked against):
//   mix/Random/phase ..... reference dsearch_desert seeds CSV (fyA, fyB) + closed-form pass fraction
//   amplitude_sub/tables/Blur/scan/sites/proxy floods ..... reference dsearch_desert CSV/JSON goldens
//   Perlin/NormalNoise/spline/classify/SurfaceGrid/desert_state/dense_flood_bridged ..... reference
//       dflood goldens at steps 1/4/8 and the MC-Measure-verified fixture cell counts
//   GPU-gate restatement (gate_pre_amp / gate_full_amp, gate.cuh:205-326): the reference has no CPU
//       form of it. Pinned by invariants instead: the merged 4-parity table equals the mean of the
//       4 per-parity tables evaluated directly in double; the full gate equals a direct 512^2
//       evaluation; blur with a delta input reproduces the weights. See tests/test_oracle_gate.py.
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>
#include <string>
#include <array>

namespace oracle {

// ------------------------------------------------------------------ RNG (standalone.hpp:13-37)
uint64_t mix(uint64_t x);
struct Rng {
    uint64_t a, b;
    static Rng seeded(uint64_t seed);          // Random(uint64_t) standalone.hpp:20-23
    uint64_t next();                           // standalone.hpp:25-29
    double uniform();                          // standalone.hpp:30
    uint32_t bounded(uint32_t n);              // standalone.hpp:31-36
};

// ------------------------------------------------------------------ Perlin (standalone.hpp:38-83)
double fade(double t);
double grad(unsigned h, double x, double y, double z);
struct Perlin {
    uint8_t perm[256];
    double ox, oy, oz, frequency, weight, fy, vy;
    int iy;
    void init(Rng r, double f, double w);
    double sample(double x, double y, double z) const;
};

// ------------------------------------------------------------------ NormalNoise (standalone.hpp:85-151)
enum Channel { Temperature, Humidity, Continentalness, Erosion, Shift, Weirdness };
struct Noise {
    Perlin oct[2][9];
    int n = 0;
    double factor = 0;
    void init(uint64_t a, uint64_t b, int channel);
    double sample(double x, double y, double z) const;
    double octaves(double x, double z, int lo, int hi) const;   // desert_proxy.hpp:14-24
};

// ------------------------------------------------------------------ climate + biome (standalone.hpp:154-239, surface.hpp)
float spline(int index, const float *v);
int classify(const int64_t *v);
struct Grid {                                  // surface.hpp:12-71
    std::vector<int32_t> cuts[5];
    std::vector<int16_t> cell;
    int dim[5] = {};
    Grid();
    size_t index(int i0, int i1, int i2, int i3, int i4) const;
    int lookup(const int64_t *q) const;
};
const Grid &grid();
struct Climate {                               // Sampler standalone.hpp:216-239 (all channels eager)
    uint64_t a, b;
    Noise noise[6];
    explicit Climate(uint64_t seed);
    void climate(int x, int y, int z, int64_t *q) const;
    void finish(double px, double pz, int y, float c, int64_t *q) const;
    int biome(int x, int z) const;             // SurfaceSampler::biome surface.hpp:80-84
    bool desert(int x, int z) const;           // surface.hpp:85-93
    int desert_state(int x, int z) const;      // surface.hpp:98-124
    float proxy_prob(int x, int z, float cut) const;    // DesertProxy::prob desert_proxy.hpp:25-30
    int proxy(int x, int z, float cut) const;           // desert_proxy.hpp:32
    int hot(int x, int z) const;               // dsearch_desert.cpp:334
};

// ------------------------------------------------------------------ flood (dense_flood.hpp:9-86)
struct Area {
    uint64_t cells = 0, sampled = 0;
    int minx = 0, maxx = 0, minz = 0, maxz = 0;
    bool edge = false, border = false;
};
// pred(x, z) -> 0 wall / 1 crossed / 2 counted. state gets the dense window (0 untouched, 1 touched,
// 2 crossed, 3 counted).
template <class Pred>
Area flood(Pred &&pred, int qx, int qz, int step, int half, bool conn8, std::vector<uint8_t> &state,
           std::vector<int> &queue);

// ------------------------------------------------------------------ seed gate
double phase_t(uint64_t seed, double *fy);                  // dsearch_desert.cpp:80-94 (== gate.cuh:71-84)
struct PerlinSide { uint8_t perm[256]; double ox, oy, oz, fy, vy; int iy; };
void init_side(uint64_t seed, int side, PerlinSide &p);     // gate.cuh:59-70
constexpr double T_W0 = 1.25 * (32.0 / 63.0) * 1.5;         // gate.cuh:47
// GPU-gate amplitudes (gate.cuh:205-326): 4-parity anti-aliased 256^2 R=4 (BOX), full 512^2 R=7,
// legacy even-parity 256^2 R=4. kern = 2R+1 unnormalised float weights, computed by the caller.
void gate_grid(const PerlinSide &p, int G, bool box, float *V);
float gate_blur_max(const float *V, int G, int R, const float *kern);
float gate_side_max(const PerlinSide &p, int G, int R, bool box, const float *kern);
float gate_amp(uint64_t seed, int G, int R, bool box, const float *kern);
float amplitude_sub(const Perlin &p, double w);             // lowfield.hpp:60-102 (CPU-default pre-gate)

// ------------------------------------------------------------------ period tables (lowfield.hpp:7-139)
constexpr int TN = 512;
struct TabIn { uint8_t perm[256]; uint8_t col[2][256]; float T[2][2][8][16]; int i0[2], k0[2]; };
void tab_setup(const Perlin &p, double w, TabIn &t);
void tabulate(const Perlin &p, double w, std::vector<float> &t);
struct Blur {
    std::vector<float> k; int r;
    explicit Blur(double sigma);
    void apply(std::vector<float> &a, std::vector<float> &t) const;
};

// ------------------------------------------------------------------ world scan (dsearch_desert.cpp:43-113, 256-308)
constexpr double BS = 337.0 / 331.0;
constexpr int WORLD = 7499996, PERIOD = 262144, NT = 60, KMIN = -30, CELL = 512;
struct TransMap {
    int16_t ib[TN][NT];
    int8_t lo[TN], hi[TN];
    explicit TransMap(double lat);
};
const TransMap &tmap_t();
const TransMap &tmap_c();
struct Site { float s, t, c, e; int x, z; };
struct Params {                                 // dsearch_desert.cpp:50-78
    double max_phase = 0.10, amp = 0.70, amp_margin = 0.01, sigma = 2.0, sigmac = 2.0;
    double wc = 0.25, we = 0.41, t = 0.55, ta = 0.25;
    int max_sites = 300;
    double pcut = 0.35, hot_min = 180;
    int multi = 1, exact_n = 2;
    double exact_min = 75, hot_pre = 0.8, csv_min = 50;
    int nms = 5000;
    double refine_mb = 70, verify_mb = 80;
    int study = 0;
    double gate_phase = 0.06, gate_amp = 0.70, gate_pthr = 0.645;
};
// Raw candidates in scan order (j, i, tz, tx), then the bucket reduction (cand) of the reference.
struct Raw { float s, t, c, e; int x, z; };
void scan_raw(const std::vector<float> tabs[6], const Params &P, std::vector<Raw> &raw);
void buckets(const std::vector<Raw> &raw, const Params &P, std::vector<Site> &cand);
void select_sites(std::vector<Site> &cand, const Params &P, std::vector<Site> &sites);

// ------------------------------------------------------------------ whole per-seed pipeline
struct Eval { size_t rank; int x, z; double proxy, s8 = -1, s4 = 0; bool edge = false; double hot = 0; };
struct SeedResult {
    uint64_t seed = 0;
    double fy[2] = {0, 0};
    bool phase_pass = false, pre_pass = false, scanned = false;
    float maxta = 0, maxtb = 0;
    size_t ncand = 0, nsites = 0;
    double best = 0;
    std::vector<Eval> ev;
    std::vector<Site> sites;
    std::vector<std::string> candidates;        // CANDIDATE lines
};
// gated = the seed came from a gate (stdin=1 / gate=1): no phase or amplitude checks
// (dsearch_desert.cpp:192-200, 216, 248).
SeedResult run_seed(uint64_t seed, const Params &P, bool gated);
std::string csv_rows(const SeedResult &r, const Params &P);   // dsearch_desert.cpp:534-542
std::string seeds_row(const SeedResult &r);                   // dsearch_desert.cpp:251, 548

}  // namespace oracle

#include "oracle_flood.inl"
