// GPU seed gate tool on top of libdsx.so, with the command line of the reference's gate tool
// (src/gpu/tgate.cu:1-7):
//   tgate START COUNT [phase=0.10] [amp=0.70] [box=1 pthr=0.645 | box=0 margin=0.01] [sig=1.0] [check=0] [calib=N]
//         [amps=1]
// Seeds are splitmix64(START + k), k in [0, COUNT). Prints "seed amp" lines (the input of
// `dsearch_desert ... stdin=1`) for the seeds passing the gate, in seed-index order.
//   calib=N  "seed pre full phase" for the first N (by seed value) phase-passed seeds of the first 2^26
//            indices (tgate.cu:53-85; the pre-gate threshold calibration), sorted by seed
//   check=1  for the first 4 passing seeds, the GPU amplitude beside the full amplitude in the CPU
//            arithmetic of tgate.cu:14-39 (cpu_ref.h gate_amp), on stderr (tgate.cu:87-91); the two use
//            different arithmetic (float corner-term tables on the GPU, double Perlin then float on the
//            CPU), so they agree to a few 1e-6, not bit for bit
//   amps=0   (this tool only) print the seeds alone, asking the library for no amplitudes, as the search
//            pipeline does: the full gate may then decide from its fast amplitude too (same seeds)
#include "../../include/dsx.h"
#include "hwq.h"
#include "../csrc/dsx_rng.h"
#include "cpu_ref.h"
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

static int fail(const char *what, int rc) {
    std::fprintf(stderr, "%s: %s %s\n", what, dsx_strerror(rc), dsx_last_error());
    return 1;
}

int main(int argc, char **argv) {
    dsx_set_hw_queues();
    if (argc < 3) {
        std::fprintf(stderr,
                     "tgate START COUNT [phase=0.10] [amp=0.70] [box=1 pthr=0.645 | box=0 margin=0.01] [check=0] [calib=N] [amps=1]\n");
        return 2;
    }
    const uint64_t start = std::stoull(argv[1]), count = std::stoull(argv[2]);
    dsx_gate_params g{0.10f, 0.70f, 0.645f, 0.01f, 1.0f, 1};   // GateCfg defaults (gate.cuh:349-355)
    int check = 0, calib = 0, want_amps = 1;
    for (int i = 3; i < argc; i++) {
        const std::string kv = argv[i];
        const auto e = kv.find('=');
        const std::string k = kv.substr(0, e);
        const double v = std::stod(kv.substr(e + 1));
        if (k == "phase") g.max_phase = float(v);
        else if (k == "amp") g.amp = float(v);
        else if (k == "margin") g.margin = float(v);
        else if (k == "check") check = int(v);
        else if (k == "calib") calib = int(v);
        else if (k == "amps") want_amps = v != 0 ? 1 : 0;
        else if (k == "box") g.box = int(v);
        else if (k == "sig") g.sig = float(v);
        else if (k == "pthr") g.pthr = float(v);
        else if (k == "tpb_pre" || k == "tpb_full" || k == "perseed") continue;   // launch shapes of the CUDA tool
        else {
            std::fprintf(stderr, "unknown key %s\n", k.c_str());
            return 2;
        }
    }
    dsx_ctx *ctx = nullptr;
    int rc = dsx_create(0, &ctx);
    if (rc) return fail("dsx_create", rc);
    const uint64_t B = 1ull << 26;   // seed indices per batch, as tgate.cu:55
    if (calib) {
        // phase survivors of the first 2^26 seed indices (count is not used, as tgate.cu:57-60), the first
        // `calib` of them by seed value, then their pre-gate and full amplitudes from the explicit list
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::vector<uint64_t>> part(nth);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; t++)
            th.emplace_back([&, t] {
                for (uint64_t k = B * t / nth; k < B * (t + 1) / nth; k++) {
                    const uint64_t seed = dsx::mix64(start + k);
                    if (dsx::temperature_phase(seed, nullptr) <= double(g.max_phase)) part[t].push_back(seed);
                }
            });
        for (auto &x : th) x.join();
        std::vector<uint64_t> seeds;
        for (auto &p : part) seeds.insert(seeds.end(), p.begin(), p.end());
        std::sort(seeds.begin(), seeds.end());
        seeds.resize(std::min<size_t>(seeds.size(), size_t(calib)));
        std::vector<float> pre(seeds.size()), full(seeds.size());
        if ((rc = dsx_gate_amps(ctx, seeds.data(), seeds.size(), &g, pre.data(), full.data())))
            return fail("dsx_gate_amps", rc);
        for (size_t i = 0; i < seeds.size(); i++)
            std::printf("%llu %.5f %.5f %.5f\n", (unsigned long long)seeds[i], pre[i], full[i],
                        dsx::temperature_phase(seeds[i], nullptr));
        dsx_destroy(ctx);
        return 0;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint64_t> seeds(1u << 20);
    std::vector<float> amps(1u << 20);
    dsx_gate_counts tot{0, 0, 0};
    int checked = 0;
    for (uint64_t pos = 0; pos < count; pos += B) {
        const uint64_t m = std::min(B, count - pos);
        uint64_t n = 0;
        dsx_gate_counts c;
        float *ap = want_amps ? amps.data() : nullptr;
        rc = dsx_gate(ctx, start + pos, m, &g, seeds.data(), ap, nullptr, seeds.size(), &n, &c);
        if (rc == DSX_E_CAPACITY) {
            seeds.resize(n);
            amps.resize(n);
            ap = want_amps ? amps.data() : nullptr;
            rc = dsx_gate(ctx, start + pos, m, &g, seeds.data(), ap, nullptr, seeds.size(), &n, &c);
        }
        if (rc) return fail("dsx_gate", rc);
        for (uint64_t i = 0; i < n; i++) {
            if (want_amps) std::printf("%llu %.5f\n", (unsigned long long)seeds[i], amps[i]);
            else std::printf("%llu\n", (unsigned long long)seeds[i]);
            if (want_amps && check && checked < 4) {
                std::fprintf(stderr, "  check seed %llu: gpu amp %.5f cpu amp %.5f\n", (unsigned long long)seeds[i],
                             amps[i], dsx::cpu::gate_amp(seeds[i], 512, 7, 2.0f));
                checked++;
            }
        }
        std::fflush(stdout);
        tot.phase += c.phase;
        tot.pre += c.pre;
        tot.full += c.full;
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "tgate: %llu seed indices, %llu phase, %llu pre-gate, %llu gated in %.2f s\n",
                 (unsigned long long)count, (unsigned long long)tot.phase, (unsigned long long)tot.pre,
                 (unsigned long long)tot.full, sec);
    dsx_destroy(ctx);
    return 0;
}
