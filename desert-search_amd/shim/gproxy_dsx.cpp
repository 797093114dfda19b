// gproxy_dsx.cpp - the reference-side drop-in for src/gpu/gproxy.cu: gproxy::{init, scan_init, scan,
// evaluate, start_gate} (src/gpu/gproxy.h:30-41) over libdsx.so's C ABI (include/dsx.h). A maintainer
// builds it with the reference's dsearch_desert.cpp and -DDRIP_GPU in place of gproxy.cu:
//   g++ -O3 -std=c++20 -DDRIP_GPU -Isrc -I<repo>/include src/dsearch_desert.cpp gproxy_dsx.cpp \
//       -L<repo>/desert-search_amd -ldsx -Wl,-rpath,<repo>/desert-search_amd -o bin/dsearch_desert_gpu
// The call sites set gproxy::t_seed = seed before :239 (gproxy::scan) and :441 (gproxy::evaluate): neither
// TabIn nor SeedNoise carries the seed, and the engine rebuilds both from it on the device.
#include "gpu/gproxy.h"
#include "dsx.h"
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <thread>
#include <vector>

namespace gproxy {
static float g_cut;
static thread_local dsx_ctx *t_ctx = nullptr;        // one context (stream + buffers) per worker thread
thread_local uint64_t t_seed;                        // set by the call sites below before scan/evaluate

static void check(int rc, const char *what) {
    if (rc != DSX_OK) { fprintf(stderr, "%s: %s %s\n", what, dsx_strerror(rc), dsx_last_error()); exit(1); }
}
static dsx_ctx *ctx() {
    if (!t_ctx) check(dsx_create(0, &t_ctx), "dsx_create");
    return t_ctx;
}

void init(int, float cut) { g_cut = cut; }           // gproxy.cu:391-401: tables are built into libdsx
void scan_init(const int16_t *, const int8_t *, const int8_t *, const int16_t *, int, const float *, int,
               const float *, int, int, int, int) {} // maps and blur weights are derived inside dsx_scan

// gproxy.h:33 - the tables come from the seed on the device, not from the host's TabIn
bool scan(const TabIn *, const ScanParams &sp, float &maxta, float &maxtb, std::vector<RawCand> &raw) {
    dsx_scan_params p{sp.ta, sp.tt, sp.wc, sp.we, 2.0, 2.0};   // Params::sigma, sigmac
    std::vector<dsx_raw_cand> c(1 << 16);
    uint64_t n = 0;
    const int rc = dsx_scan(ctx(), &t_seed, 1, &p, &maxta, &maxtb, c.data(), c.size(), &n);
    if (rc == DSX_E_CAPACITY) return false;          // "gscan: candidate overflow ... CPU fallback"
    check(rc, "dsx_scan");
    raw.resize(n);
    for (uint64_t i = 0; i < n; i++) raw[i] = {c[i].s, c[i].t, c[i].c, c[i].e, c[i].x, c[i].z};
    return true;                                     // raw is in the CPU's scan order
}

// gproxy.h:37 - zone cell (i, j) is the quart (x0 + 32 i, z0 + 32 j), rows z-major. The FP64
// DesertProxy of desert_proxy.hpp:14-32 (the CPU proxy's result), not the FP32 k_proxy.
void evaluate(const SeedNoise &, const std::vector<Zone> &zones, std::vector<std::vector<uint8_t>> &bits) {
    bits.resize(zones.size());
    for (size_t k = 0; k < zones.size(); k++) {
        const Zone &z = zones[k];
        std::vector<int32_t> xs(size_t(z.nx) * z.nz), zs(xs.size());
        for (int j = 0; j < z.nz; j++)
            for (int i = 0; i < z.nx; i++) { xs[size_t(j) * z.nx + i] = z.x0 + 32 * i; zs[size_t(j) * z.nx + i] = z.z0 + 32 * j; }
        std::vector<int8_t> out(xs.size());
        check(dsx_predicates(ctx(), t_seed, DSX_PRED_PROXY, g_cut, xs.data(), zs.data(), xs.size(), out.data()),
              "dsx_predicates");
        bits[k].resize(out.size());
        for (size_t i = 0; i < out.size(); i++) bits[k][i] = out[i] == 2;
    }
}

// gproxy.h:40-41
void start_gate(int device, uint64_t start, uint64_t count, float phase, float amp, float pthr,
                std::function<void(uint64_t)> push, std::function<void()> done) {
    std::thread([=] {
        dsx_ctx *g = nullptr;
        check(dsx_create(device, &g), "dsx_create");
        dsx_gate_params p{phase, amp, pthr, 0.01f, 1.0f, 1};
        std::vector<uint64_t> seeds(1 << 16);
        const uint64_t B = 1ull << 26;               // same batch as gate.cuh:363
        for (uint64_t base = 0; base < count; base += B) {
            uint64_t n = 0;
            dsx_gate_counts c;
            int rc = dsx_gate(g, start + base, std::min(B, count - base), &p, seeds.data(), nullptr, nullptr,
                              seeds.size(), &n, &c);
            if (rc == DSX_E_CAPACITY) {              // n holds the true count: grow and redo the batch
                seeds.resize(n);
                rc = dsx_gate(g, start + base, std::min(B, count - base), &p, seeds.data(), nullptr, nullptr,
                              seeds.size(), &n, &c);
            }
            check(rc, "dsx_gate");
            for (uint64_t i = 0; i < n; i++) push(seeds[i]);   // sorted by seed index
        }
        dsx_destroy(g);
        done();
    }).detach();
}
}  // namespace gproxy
