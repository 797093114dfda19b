// Whole-world desert searcher on MI355X: the reference's command line on top of libdsx.so.
//   dsearch_desert START COUNT OUT_PREFIX [THREADS] [key=value ...]     seeds = splitmix64(START + k)
// Same keys, defaults (struct Params, dsearch_desert.cpp:50-78), checks, outputs (PREFIX.csv,
// PREFIX.seeds.csv, PREFIX.json) and stderr lines (CANDIDATE, best, progress, done) as the reference.
// Every stage runs on the GPU (gate, tables, scan, hot zones, proxy and exact floods); the gpu and gscan
// keys are accepted for compatibility. gate=1 runs the GPU seed gate over [START, START+COUNT) (COUNT 0 =
// unbounded); stdin=1 reads gated seeds ("seed [amp]" lines) from stdin; otherwise the CPU-default gate
// semantics (phase <= `phase`, amplitude_sub >= amp - amp_margin, maxta + maxtb >= amp) apply.
// gscancheck=N and gpucheck=N (with gpu=1) compare the first N seeds' tables and scan, and the first N big
// hot zones' proxy predicates, with the reference's CPU arithmetic and print the reference's lines.
// SIGINT / SIGTERM stop the search as the reference's handler does: seeds in flight are finished and
// written, then the JSON is written (dsearch_desert.cpp:40-41, 586-598).
#include "../../include/dsx.h"
#include "hwq.h"
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static volatile sig_atomic_t interrupted = 0;
static void on_signal(int) {
    interrupted = 1;
    dsx_interrupt(1);
}

int main(int argc, char **argv) {
    dsx_set_hw_queues();
    if (argc < 4) {
        std::fprintf(stderr, "dsearch_desert START COUNT OUT_PREFIX [threads] [key=value...]\n");
        return 2;
    }
    const uint64_t start = std::stoull(argv[1]), count = std::stoull(argv[2]);
    const std::string prefix = argv[3];
    const int threads = argc > 4 ? std::stoi(argv[4]) : 11;
    dsx_search_params P;
    dsx_search_defaults(&P);
    P.workers = threads < 1 ? 1 : (threads > 64 ? 64 : threads);   // THREADS -> pipeline workers
    for (int i = 5; i < argc; i++) {
        const std::string kv = argv[i];
        const auto eq = kv.find('=');
        const std::string k = kv.substr(0, eq);
        const double v = std::stod(kv.substr(eq + 1));
        struct { const char *name; double *d; int32_t *n; } keys[] = {
            {"gpu", nullptr, &P.gpu}, {"gpucheck", nullptr, &P.gpucheck}, {"stdin", nullptr, &P.from_stdin},
            {"study", nullptr, &P.study}, {"phase", &P.max_phase, nullptr}, {"amp", &P.amp, nullptr},
            {"amp_margin", &P.amp_margin, nullptr}, {"sigma", &P.sigma, nullptr}, {"sigmac", &P.sigmac, nullptr},
            {"wc", &P.wc, nullptr}, {"we", &P.we, nullptr}, {"t", &P.t, nullptr}, {"ta", &P.ta, nullptr},
            {"sites", nullptr, &P.max_sites}, {"nms", nullptr, &P.nms}, {"pcut", &P.pcut, nullptr},
            {"multi", nullptr, &P.multi}, {"hot_min", &P.hot_min, nullptr}, {"exact_n", nullptr, &P.exact_n},
            {"exact_min", &P.exact_min, nullptr}, {"csv_min", &P.csv_min, nullptr}, {"hot_pre", &P.hot_pre, nullptr},
            {"gate", nullptr, &P.gate}, {"gscan", nullptr, &P.gscan}, {"gscancheck", nullptr, &P.gscancheck},
            {"gate_phase", &P.gate_phase, nullptr}, {"gate_amp", &P.gate_amp, nullptr},
            {"gate_pthr", &P.gate_pthr, nullptr}, {"refine", &P.refine_mb, nullptr}, {"verify", &P.verify_mb, nullptr},
            {"prepass", nullptr, &P.prepass}, {"mcverify", nullptr, &P.mcverify}};
        bool ok = false;
        for (auto &e : keys)
            if (k == e.name) {
                if (e.d) *e.d = v;
                else *e.n = int32_t(v);
                ok = true;
                break;
            }
        if (!ok) {
            std::fprintf(stderr, "unknown key %s\n", k.c_str());
            return 2;
        }
    }
    if ((P.gscan || P.gscancheck) && !P.gpu) { std::fprintf(stderr, "gscan=1 needs gpu=1\n"); return 2; }
    if (P.gate) {
        if (!P.gpu) { std::fprintf(stderr, "gate=1 needs gpu=1\n"); return 2; }
        P.from_stdin = 1;
    }
    FILE *out = std::fopen((prefix + ".csv").c_str(), "wx");
    FILE *sout = std::fopen((prefix + ".seeds.csv").c_str(), "wx");
    if (!out || !sout) {
        std::fprintf(stderr, "cannot create %s outputs (exist?)\n", prefix.c_str());
        return 2;
    }
    std::fprintf(out, "seed,fyA,fyB,maxta,maxtb,rank,x,z,score,t0,c0,e0,proxy,s8,s4,edge,hot\n");
    std::fprintf(sout, "seed,fyA,fyB,maxta,maxtb,candidates,sites,best\n");
    std::fclose(out);
    std::fclose(sout);
    if (P.study) {   // dsearch_desert.cpp:147-150
        FILE *zout = std::fopen((prefix + ".zones.csv").c_str(), "wx");
        if (!zout) {
            std::fprintf(stderr, "cannot create %s outputs (exist?)\n", prefix.c_str());
            return 2;
        }
        std::fprintf(zout, "seed,phT,amp,phE0,phE1,phC0,phC1,rank,x,z,t0,c0,e0,hot,proxy,f_ocean,f_badl,f_mang,f_good,mean_c,"
                           "mean_e,sd_e\n");
        std::fclose(zout);
    }
    std::signal(SIGINT, on_signal);
    std::signal(SIGTERM, on_signal);

    dsx_ctx *ctx = nullptr;
    int rc = dsx_create(0, &ctx);
    if (rc) {
        std::fprintf(stderr, "dsx_create: %s %s\n", dsx_strerror(rc), dsx_last_error());
        return 1;
    }
    dsx_search_counts C{};
    const auto begin = std::chrono::steady_clock::now();
    auto last = begin;
    auto progress = [&](bool force) {
        const auto now = std::chrono::steady_clock::now();
        if (!force && std::chrono::duration<double>(now - last).count() <= 30) return;
        last = now;
        // dsearch_desert.cpp:552-554; core-s are the pipeline workers' seconds per stage (dsx_search_counts)
        std::fprintf(stderr,
                     "%llu seeds, %llu passed, %llu scanned, %llu sites; %.2f passed/s; core-s tables %.0f scan %.0f "
                     "sites %.0f (proxy %.0f hot %.0f gpu %.0f); hot zones %llu, big %llu\n",
                     (unsigned long long)C.done, (unsigned long long)C.passed, (unsigned long long)C.scanned,
                     (unsigned long long)C.sites, C.passed / std::chrono::duration<double>(now - begin).count(),
                     C.sec_tables, C.sec_scan, C.sec_sites, C.sec_proxy, C.sec_hot, C.sec_gpu,
                     (unsigned long long)C.hot_zones, (unsigned long long)C.big_zones);
    };
    if (P.from_stdin && !P.gate) {
        char line[256];
        std::vector<uint64_t> seeds;
        auto flush = [&]() {
            if (seeds.empty()) return 0;
            const int r = dsx_search(ctx, &P, DSX_SEARCH_LIST, 0, 0, seeds.data(), seeds.size(), prefix.c_str(), &C);
            seeds.clear();
            progress(false);
            return r;
        };
        while (!interrupted && rc == 0 && std::fgets(line, sizeof line, stdin)) {
            seeds.push_back(std::strtoull(line, nullptr, 10));
            if (seeds.size() >= 4096) rc = flush();
        }
        if (!rc) rc = flush();
    } else {
        const uint64_t total = P.gate ? (count ? count : 100000000000000ULL) : count;
        const uint64_t step = P.gate ? (1ull << 26) : (1ull << 24);
        for (uint64_t pos = 0; pos < total && !interrupted && rc == 0; pos += step) {
            const uint64_t m = total - pos < step ? total - pos : step;
            rc = dsx_search(ctx, &P, P.gate ? DSX_SEARCH_GATE : DSX_SEARCH_ENUM, start + pos, m, nullptr, 0, prefix.c_str(),
                            &C);
            progress(false);
        }
    }
    if (rc == DSX_E_INTERRUPTED) rc = 0;   // the reference leaves with status 0 after an interrupt
    if (rc) std::fprintf(stderr, "dsx_search: %s %s\n", dsx_strerror(rc), dsx_last_error());
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
    FILE *m = std::fopen((prefix + ".json").c_str(), "w");
    std::fprintf(m,
                 "{\"start\":%llu,\"done\":%llu,\"passed\":%llu,\"scanned\":%llu,\"amp\":%g,\"sites\":%llu,\"seconds\":%.3f,"
                 "\"threads\":%d,\"phase\":%g,\"sigma\":%g,\"wc\":%g,\"we\":%g,\"t\":%g,\"ta\":%g,\"max_sites\":%d,\"nms\":%d,"
                 "\"cpu_tables\":%.3f,\"cpu_scan\":%.3f,\"cpu_sites\":%.3f}\n",
                 (unsigned long long)start, (unsigned long long)C.done, (unsigned long long)C.passed,
                 (unsigned long long)C.scanned, P.amp, (unsigned long long)C.sites, sec, threads, P.max_phase, P.sigma, P.wc,
                 P.we, P.t, P.ta, P.max_sites, P.nms, C.sec_tables, C.sec_scan, C.sec_sites);
    std::fclose(m);
    std::fprintf(stderr, "done: %llu seeds (%llu passed) in %.1fs\n", (unsigned long long)C.done,
                 (unsigned long long)C.passed, sec);
    dsx_destroy(ctx);
    return rc ? 1 : 0;
}
