// C ABI of libdsx.so (include/dsx.h): context management and the seed-gate host loop.
#include "ctx.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace dsx {
thread_local std::string g_last_error;
int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[320];
    std::snprintf(buf, sizeof buf, "%s: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_last_error = buf;
    return DSX_E_HIP;
}
}  // namespace dsx
using dsx::g_last_error;

// Gate blur weights: expf(-0.5 d^2 / sig^2) in float, as gate.cuh:216 forms them, evaluated here
// with the host libm; ksum is their float sum in index order and norm = 1 / (ksum * ksum).
dsx::GateKernel dsx::gate_kernel(int R, float sig) {
    dsx::GateKernel K{};
    for (int t = 0; t <= 2 * R; t++) {
        const float d = float(t - R);
        K.k[t] = expf(-0.5f * d * d / (sig * sig));
    }
    float ks = 0;
    for (int t = 0; t <= 2 * R; t++) ks += K.k[t];
    K.norm = 1.f / (ks * ks);
    K.wd = double(float(1.25 * (32.0 / 63.0) * 1.5));
    // the normalised 2D weights the truncated blur (taps 1 .. 2R - 1) leaves out, in double, with a margin
    // of 1e-3 for this evaluation's own roundings; the weights are positive
    if (R >= 2) {
        double s = 0, st = 0;
        for (int t = 0; t <= 2 * R; t++) {
            s += K.k[t];
            if (t > 0 && t < 2 * R) st += K.k[t];
        }
        K.trunc = std::nextafter(float((s * s - st * st) * double(K.norm) * 1.001), INFINITY);
    }
    // integer weights of the 16-bit fast amplitudes (R = 4: the truncated taps 1 .. 7 at SK = 2^14; R = 7: all
    // 15 taps at SK = 2^13, so that the sums of 15 products stay below 2^31): rounding them moves a pass's
    // output by at most sum_d |kq_d / SK - w_d| times the largest |input|, so the normalised output by that
    // times the other pass's taps' sum times norm, per pass and per unit of |grid value|; the horizontal sums,
    // rounded to units of 2^16 (2^15 / (SK S) = 2^15 / (32000 SK) b of a grid value each, S = 32000 / b), move
    // it by that times the vertical taps' sum times norm. wq is the sum of the three, with 1e-3 for this
    // evaluation's roundings and for the fixed-point grid's own excess over the grid bound.
    if (R == 4 || R == 7) {
        const int t0 = R == 4 ? 1 : 0, W = R == 4 ? 7 : 15, NP = (W + 1) / 2;
        const double SK = R == 4 ? 16384.0 : 8192.0;
        int kq[16] = {0};
        double dk = 0, st = 0;
        for (int d = 0; d < W; d++) {
            kq[d] = int(std::lrint(double(K.k[t0 + d]) * SK));
            dk += std::fabs(kq[d] / SK - double(K.k[t0 + d]));
            st += K.k[t0 + d];
        }
        const auto w = [&](int d) { return d < 0 || d >= W ? 0 : kq[d]; };
        const auto pk = [](int lo, int hi) { return (uint32_t(lo) & 0xffffu) | (uint32_t(hi) << 16); };
        for (int m = 0; m < NP; m++) {
            K.kq2[m] = pk(w(2 * m - 1), w(2 * m));
            K.kq2[NP + m] = pk(w(2 * m), w(2 * m + 1));
        }
        K.wq = std::nextafter(float((2.0 * dk + 32768.0 / (32000.0 * SK)) * st * double(K.norm) * 1.001), INFINITY);
    }
    return K;
}
bool dsx::valid_gate_params(const dsx_gate_params *p) {
    return p && p->box >= 0 && p->box <= 1 && std::isfinite(p->max_phase) && std::isfinite(p->amp) &&
           std::isfinite(p->pthr) && std::isfinite(p->margin) && p->sig > 0 && std::isfinite(p->sig);
}

namespace {
using dsx::gate_kernel;
using dsx::valid_gate_params;
// CPU-default pre-gate kernel (lowfield.hpp:92-96): r = 4, float(exp(-d^2/2)) normalised by their
// double sum; octave weight factor * weight of temperature octave 0 (dsearch_desert.cpp:217).
dsx::GateKernel cpu_gate_kernel() {
    dsx::GateKernel K{};
    constexpr int r = 4;
    double s = 0;
    for (int d = -r; d <= r; d++) {
        K.k[d + r] = float(std::exp(-0.5 * d * d));
        s += K.k[d + r];
    }
    for (int d = 0; d <= 2 * r; d++) K.k[d] = float(K.k[d] / s);
    K.norm = 1.f;
    const dsx_noise_cfg &c = DSX_NOISE[DSX_CH_T];
    const double w0 = double(1 << (c.len - 1)) / ((1 << c.len) - 1);
    int lo = 0, hi = c.len - 1;
    while (c.amp[lo] == 0) lo++;
    while (c.amp[hi] == 0) hi--;
    const double factor = (5. * (hi - lo + 1)) / (3. * (hi - lo + 2));
    K.wd = factor * (w0 * c.amp[0]);
    return K;
}
// Settings chosen by the autotuner for this device: tuned.json beside the library holds an "env" object of
// "DSX_*": "value" pairs, each of which becomes the default of that environment variable (a variable that
// is already set wins; DSX_TUNED=0 skips the file). Applied once, at the first dsx_create and before its
// first HIP call, so every setting the library reads later (launch shapes, workers, batching) sees it; the
// settings are all result-invariant. A malformed file is ignored with a message on stderr.
void apply_tuned() {
    static std::once_flag once;
    std::call_once(once, [] {
        const char *off = std::getenv("DSX_TUNED");
        if (off && *off == '0') return;
        Dl_info info;
        if (!dladdr(reinterpret_cast<void *>(&apply_tuned), &info) || !info.dli_fname) return;
        std::string path(info.dli_fname);
        path = path.substr(0, path.rfind('/') + 1) + "tuned.json";
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) return;
        std::string txt;
        char buf[4096];
        for (size_t n; (n = std::fread(buf, 1, sizeof buf, f)) > 0;) txt.append(buf, n);
        std::fclose(f);
        // the "env" object: a flat list of string keys and string or number values
        size_t p = txt.find("\"env\"");
        p = p == std::string::npos ? p : txt.find('{', p);
        const size_t e = p == std::string::npos ? p : txt.find('}', p);
        if (e == std::string::npos) {
            std::fprintf(stderr, "dsx: %s has no \"env\" object; ignored\n", path.c_str());
            return;
        }
        std::vector<std::pair<std::string, std::string>> kv;
        const auto parse = [&] {
            for (size_t i = p + 1; i < e;) {
                const size_t k0 = txt.find('"', i);
                if (k0 == std::string::npos || k0 >= e) return true;
                const size_t k1 = txt.find('"', k0 + 1), c = k1 == std::string::npos ? k1 : txt.find(':', k1);
                if (c == std::string::npos || c >= e) return false;
                const size_t v0 = txt.find_first_not_of(" \t\r\n", c + 1);
                if (v0 == std::string::npos || v0 >= e) return false;
                const bool quoted = txt[v0] == '"';
                const size_t v1 = quoted ? txt.find('"', v0 + 1) : txt.find_first_of(",} \t\r\n", v0);
                if (v1 == std::string::npos || v1 > e || (quoted && v1 == e)) return false;
                kv.emplace_back(txt.substr(k0 + 1, k1 - k0 - 1), txt.substr(v0 + quoted, v1 - v0 - quoted));
                i = v1 + quoted;
            }
            return true;
        };
        if (!parse()) {
            std::fprintf(stderr, "dsx: malformed \"env\" object in %s; ignored\n", path.c_str());
            return;
        }
        for (auto &[k, v] : kv)
            if (k.rfind("DSX_", 0) == 0 && k != "DSX_TUNED") setenv(k.c_str(), v.c_str(), 0);
    });
}
}  // namespace

static int ctx_alloc(dsx_ctx *c);

int dsx_ctx_device(const dsx_ctx *c) { return c->device; }

extern "C" {

const char *dsx_strerror(int code) {
    switch (code) {
    case DSX_OK: return "ok";
    case DSX_E_ARG: return "invalid argument";
    case DSX_E_HIP: return "HIP runtime error";
    case DSX_E_CAPACITY: return "capacity exceeded";
    case DSX_E_NODEVICE: return "no usable gfx950 device";
    case DSX_E_BUSY: return "request not finished";
    case DSX_E_INTERRUPTED: return "interrupted";
    default: return "unknown error";
    }
}
const char *dsx_last_error(void) { return g_last_error.c_str(); }

int dsx_create(int device, dsx_ctx **out) {
    if (!out) return DSX_E_ARG;
    *out = nullptr;
    apply_tuned();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        g_last_error = "no HIP device";
        return DSX_E_NODEVICE;
    }
    HIPCK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_last_error = std::string("device is ") + prop.gcnArchName + ", library built for gfx950";
        return DSX_E_NODEVICE;
    }
    // host wait mode of stream synchronisations (DSX_SYNC=spin|yield|block; default: the runtime's)
    if (const char *m = std::getenv("DSX_SYNC")) {
        const unsigned f = !std::strcmp(m, "spin") ? hipDeviceScheduleSpin
                           : !std::strcmp(m, "yield") ? hipDeviceScheduleYield
                           : !std::strcmp(m, "block") ? hipDeviceScheduleBlockingSync : hipDeviceScheduleAuto;
        (void)hipSetDeviceFlags(f);
    }
    dsx_ctx *c = new dsx_ctx();
    c->device = device;
    const int rc = ctx_alloc(c);
    if (rc != DSX_OK) {
        dsx_destroy(c);
        return rc;
    }
    *out = c;
    return DSX_OK;
}

static int ctx_alloc(dsx_ctx *c) {
    HIPCK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    c->cap1 = 1u << 18;
    const size_t n1 = c->cap1;
    HIPCK(hipMalloc(&c->d_k1, n1 * 4));
    HIPCK(hipMalloc(&c->d_i2, n1 * 4));
    HIPCK(hipMalloc(&c->d_k3, n1 * 4));
    HIPCK(hipMalloc(&c->d_a3, n1 * 4));
    HIPCK(hipMalloc(&c->d_all, n1 * 4));
    HIPCK(hipMalloc(&c->d_cnt, dsx::GATE_NCNT * 4));
    HIPCK(hipMalloc(&c->d_sides, 2 * n1 * sizeof(dsx::SideParams)));
    HIPCK(hipMalloc(&c->d_seeds, n1 * 8));
    HIPCK(hipHostMalloc(&c->h_cnt, 256, hipHostMallocDefault));
    HIPCK(hipHostMalloc(&c->h_k3, n1 * 4, hipHostMallocDefault));
    HIPCK(hipHostMalloc(&c->h_a3, n1 * 4, hipHostMallocDefault));
    for (auto &e : c->ev) HIPCK(hipEventCreate(&e));
    return DSX_OK;
}

int dsx_destroy(dsx_ctx *c) {
    if (!c) return DSX_E_ARG;
    dsx_ctx_free_workers(c);
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    dsx_ctx_free_ext(c);
    for (void *p : {(void *)c->d_k1, (void *)c->d_i2, (void *)c->d_k3, (void *)c->d_a3, (void *)c->d_all, (void *)c->d_cnt,
                    (void *)c->d_sides, (void *)c->d_seeds, (void *)c->d_tsp, (void *)c->d_tabs, (void *)c->d_bandmax,
                    (void *)c->d_act, (void *)c->d_ncand, (void *)c->d_cand, (void *)c->d_ibT, (void *)c->d_ibC, (void *)c->d_band,
                    (void *)c->d_loT, (void *)c->d_hiT})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)c->h_cnt, (void *)c->h_k3, (void *)c->h_a3, (void *)c->h_seeds, (void *)c->h_smax,
                    (void *)c->h_sn, (void *)c->h_cand})
        if (p) (void)hipHostFree(p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    dsx::gate_graph_free(c->ggraph);
    if (c->h_gout) (void)hipHostFree(c->h_gout);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return DSX_OK;
}

int dsx_set_gate_graph(dsx_ctx *c, int on) {
    if (!c) return DSX_E_ARG;
    c->gate_graph = on ? 1 : 0;
    return DSX_OK;
}

}  // extern "C"

namespace {
using Pass = std::vector<std::pair<uint64_t, float>>;

// The gate over [start + off, start + off + count) with a host round trip per stage (survivor counts read
// back before each stage's launch); survivors appended as (off + k, amplitude). Per-kernel times in prof_*.
int gate_eager(dsx_ctx *c, uint64_t start, uint64_t off, uint64_t count, const dsx_gate_params *p, Pass &pass,
               uint64_t tot[3]) {
    const dsx::GateKernel kpre = gate_kernel(4, p->box ? p->sig : 1.0f);
    const dsx::GateKernel kfull = gate_kernel(7, 2.0f);
    const double thr_pre = p->box ? double(p->pthr) : double(p->amp - p->margin);
    const int mode_pre = p->box ? 0 : 2;
    const double maxph = double(p->max_phase);
    uint64_t B = 1ull << 24;   // seed indices per batch (expected phase survivors ~0.72% << cap1)
    for (uint64_t base = off; base < off + count;) {
        const uint32_t nb = uint32_t(std::min<uint64_t>(B, off + count - base));
        HIPCK(hipMemsetAsync(c->d_cnt, 0, 16, c->st));
        HIPCK(hipEventRecord(c->ev[0], c->st));
        dsx::launch_phase(c->st, start + base, nb, maxph, c->d_k1, c->d_cnt, c->cap1);
        HIPCK(hipGetLastError());
        HIPCK(hipEventRecord(c->ev[1], c->st));
        HIPCK(hipMemcpyAsync(c->h_cnt, c->d_cnt, 4, hipMemcpyDeviceToHost, c->st));
        HIPCK(hipStreamSynchronize(c->st));
        const uint32_t n1 = c->h_cnt[0];
        if (n1 > c->cap1) {                // phase survivors overflow the batch buffers: smaller batch
            B = std::max<uint64_t>(1, B / 2);
            continue;
        }
        float ms;
        HIPCK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->prof_ms[0] += ms;
        c->prof_work[0] += nb;
        uint32_t n2 = 0, n3 = 0;
        if (n1) {
            HIPCK(hipEventRecord(c->ev[2], c->st));
            dsx::launch_init(c->st, start + base, c->d_k1, nullptr, n1, c->d_sides);
            HIPCK(hipGetLastError());
            HIPCK(hipEventRecord(c->ev[3], c->st));
            dsx::launch_amp(c->st, mode_pre, dsx::GATE_EXACT, c->d_sides, nullptr, n1, kpre, thr_pre, nullptr, c->d_i2, nullptr,
                            c->d_cnt + 1, c->cap1, nullptr);
            HIPCK(hipGetLastError());
            HIPCK(hipEventRecord(c->ev[4], c->st));
            HIPCK(hipMemcpyAsync(c->h_cnt + 1, c->d_cnt + 1, 4, hipMemcpyDeviceToHost, c->st));
            HIPCK(hipStreamSynchronize(c->st));
            n2 = std::min(c->h_cnt[1], c->cap1);
            HIPCK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
            c->prof_ms[1] += ms;
            c->prof_work[1] += 2ull * n1;
            HIPCK(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
            c->prof_ms[2] += ms;
            c->prof_work[2] += n1;
            if (n2) {
                HIPCK(hipEventRecord(c->ev[5], c->st));
                dsx::launch_amp(c->st, 1, dsx::GATE_EXACT, c->d_sides, c->d_i2, n2, kfull, double(p->amp), c->d_k1, c->d_k3, c->d_a3, c->d_cnt + 2,
                                c->cap1, nullptr);
                HIPCK(hipGetLastError());
                HIPCK(hipEventRecord(c->ev[6], c->st));
                HIPCK(hipMemcpyAsync(c->h_cnt + 2, c->d_cnt + 2, 4, hipMemcpyDeviceToHost, c->st));
                HIPCK(hipStreamSynchronize(c->st));
                n3 = std::min(c->h_cnt[2], c->cap1);
                HIPCK(hipEventElapsedTime(&ms, c->ev[5], c->ev[6]));
                c->prof_ms[3] += ms;
                c->prof_work[3] += n2;
                if (n3) {
                    HIPCK(hipMemcpyAsync(c->h_k3, c->d_k3, size_t(n3) * 4, hipMemcpyDeviceToHost, c->st));
                    HIPCK(hipMemcpyAsync(c->h_a3, c->d_a3, size_t(n3) * 4, hipMemcpyDeviceToHost, c->st));
                    HIPCK(hipStreamSynchronize(c->st));
                    for (uint32_t i = 0; i < n3; i++) pass.push_back({base + c->h_k3[i], c->h_a3[i]});
                }
            }
        }
        tot[0] += n1;
        tot[1] += n2;
        tot[2] += n3;
        base += nb;
    }
    return DSX_OK;
}

// Gate graph configuration: chunks of nmax indices, the largest power of two up to 2^26 whose expected
// phase survivors 2 s^2 nmax (s = max_phase <= 0.5), with a quarter and 4096 to spare, fit the survivor
// buffers (cap1), shrunk to the smallest power of two that holds the whole call; grid = that survivor
// bound. A chunk whose survivors exceed the grid is redone by gate_eager.
void gate_graph_shape(const dsx_ctx *c, double maxph, uint64_t count, uint32_t &nmax, uint32_t &grid) {
    const double s = std::min(std::max(maxph, 0.0), 0.5), f = 2.0 * s * s;
    uint64_t n = 1ull << 26;
    while (n > 256 && 1.25 * f * double(n) + 4096 > double(c->cap1)) n >>= 1;
    while (n > 256 && n / 2 >= count) n >>= 1;
    nmax = uint32_t(n);
    const uint64_t g = (uint64_t(1.25 * f * double(n)) + 4096 + 255) & ~uint64_t(255);
    grid = uint32_t(std::min<uint64_t>(g, c->cap1));
}

// want_amps false: the caller takes no amplitudes, so the full gate may decide from its fast amplitude too
// (DSX_GATE_FAST_FULL=0 keeps the reference's arithmetic for every full-gate amplitude).
int gate_graph_range(dsx_ctx *c, uint64_t start, uint64_t count, const dsx_gate_params *p, Pass &pass, uint64_t tot[3],
                     bool want_amps) {
    static const bool fast_full_on = [] {
        const char *e = std::getenv("DSX_GATE_FAST_FULL");
        return !(e && *e == '0');
    }();
    dsx::GateStages S;
    std::memset(&S, 0, sizeof S);   // compared bytewise with the graph's configuration (padding included)
    S.full_fast = (!want_amps && fast_full_on) ? 1 : 0;
    S.maxph = double(p->max_phase);
    S.thr_pre = p->box ? double(p->pthr) : double(p->amp - p->margin);
    S.amp = double(p->amp);
    S.mode_pre = p->box ? 0 : 2;
    S.kpre = gate_kernel(4, p->box ? p->sig : 1.0f);
    S.kfull = gate_kernel(7, 2.0f);
    uint32_t nmax, grid;
    gate_graph_shape(c, S.maxph, count, nmax, grid);
    if (!c->h_gout) HIPCK(hipHostMalloc(&c->h_gout, (4 + size_t(c->cap1)) * 8 + size_t(c->cap1) * 4, hipHostMallocDefault));
    uint64_t *h_counts = c->h_gout, *h_kidx = c->h_gout + 4;
    float *h_amps = reinterpret_cast<float *>(h_kidx + c->cap1);
    dsx::GateGraph &G = c->ggraph;
    const bool same = G.exec && G.C.grid == grid && G.nmax == nmax && std::memcmp(&G.S, &S, sizeof S) == 0;
    if (!same) {
        // the fast chain: decisions from the fast amplitude where provably equal, the reference's arithmetic
        // for the rest and for every amplitude written
        if (!c->d_band) HIPCK(hipMalloc(&c->d_band, size_t(c->cap1) * 16 * sizeof(float)));
        const dsx::GateChain C{c->d_cnt, c->d_k1, c->d_i2, c->d_k3, c->d_a3, c->d_sides, grid, true,
                               reinterpret_cast<uint32_t *>(c->d_all), c->d_band};
        HIPCK(dsx::gate_graph_build(G, C, S, nmax, dsx::GateOut{nullptr, h_amps, h_kidx, c->cap1, h_counts}));
    }
    for (uint64_t base = 0; base < count; base += nmax) {
        const uint32_t m = uint32_t(std::min<uint64_t>(nmax, count - base));
        HIPCK(dsx::gate_graph_launch(G, c->st, start + base, m));
        HIPCK(hipStreamSynchronize(c->st));
        if (h_counts[0] > grid) {   // more phase survivors than the graph holds: this chunk the eager way
            if (const int rc = gate_eager(c, start, base, m, p, pass, tot)) return rc;
            continue;
        }
        const uint64_t n3 = h_counts[3];
        for (uint64_t i = 0; i < n3; i++) pass.push_back({base + h_kidx[i], h_amps[i]});
        tot[0] += h_counts[0];
        tot[1] += h_counts[1];
        tot[2] += h_counts[2];
    }
    return DSX_OK;
}
}  // namespace

extern "C" {

int dsx_gate(dsx_ctx *c, uint64_t start, uint64_t count, const dsx_gate_params *p, uint64_t *seeds, float *amps,
             uint64_t *kidx, uint64_t cap, uint64_t *n_out, dsx_gate_counts *counts) {
    if (!c || !valid_gate_params(p) || !n_out) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    if (c->gate_graph < 0) {
        const char *e = std::getenv("DSX_GATE_GRAPH");
        c->gate_graph = (e && *e == '0') ? 0 : 1;
    }
    Pass pass;
    uint64_t tot[3] = {0, 0, 0};
    for (int i = 0; i < 4; i++) { c->prof_ms[i] = 0; c->prof_work[i] = 0; }
    const int rc = c->gate_graph ? gate_graph_range(c, start, count, p, pass, tot, amps != nullptr)
                                 : gate_eager(c, start, 0, count, p, pass, tot);
    if (rc) return rc;
    std::sort(pass.begin(), pass.end());
    if (counts) { counts->phase = tot[0]; counts->pre = tot[1]; counts->full = tot[2]; }
    *n_out = pass.size();
    const uint64_t nw = std::min<uint64_t>(cap, pass.size());
    for (uint64_t i = 0; i < nw; i++) {
        if (seeds) seeds[i] = dsx::mix64(start + pass[i].first);
        if (amps) amps[i] = pass[i].second;
        if (kidx) kidx[i] = pass[i].first;
    }
    return pass.size() > cap ? DSX_E_CAPACITY : DSX_OK;
}

int dsx_gate_cpu(dsx_ctx *c, uint64_t start, uint64_t count, double max_phase, double amp, double margin,
                 uint64_t *seeds, float *est, uint64_t *kidx, uint64_t cap, uint64_t *n_out, uint64_t *n_phase) {
    if (!c || !n_out || !std::isfinite(max_phase) || !std::isfinite(amp) || !std::isfinite(margin)) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    const dsx::GateKernel K = cpu_gate_kernel();
    const double thr = amp - margin;
    std::vector<std::pair<uint64_t, float>> pass;
    uint64_t nph = 0;
    uint64_t B = 1ull << 24;
    for (uint64_t base = 0; base < count;) {
        const uint32_t nb = uint32_t(std::min<uint64_t>(B, count - base));
        HIPCK(hipMemsetAsync(c->d_cnt, 0, 16, c->st));
        dsx::launch_phase(c->st, start + base, nb, max_phase, c->d_k1, c->d_cnt, c->cap1);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(c->h_cnt, c->d_cnt, 4, hipMemcpyDeviceToHost, c->st));
        HIPCK(hipStreamSynchronize(c->st));
        const uint32_t n1 = c->h_cnt[0];
        if (n1 > c->cap1) {
            B = std::max<uint64_t>(1, B / 2);
            continue;
        }
        if (n1) {
            dsx::launch_init(c->st, start + base, c->d_k1, nullptr, n1, c->d_sides);
            dsx::launch_amp(c->st, 2, dsx::GATE_EXACT, c->d_sides, nullptr, n1, K, thr, c->d_k1, c->d_k3, c->d_a3, c->d_cnt + 1, c->cap1,
                            nullptr);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(c->h_cnt + 1, c->d_cnt + 1, 4, hipMemcpyDeviceToHost, c->st));
            HIPCK(hipStreamSynchronize(c->st));
            const uint32_t n2 = std::min(c->h_cnt[1], c->cap1);
            if (n2) {
                HIPCK(hipMemcpyAsync(c->h_k3, c->d_k3, size_t(n2) * 4, hipMemcpyDeviceToHost, c->st));
                HIPCK(hipMemcpyAsync(c->h_a3, c->d_a3, size_t(n2) * 4, hipMemcpyDeviceToHost, c->st));
                HIPCK(hipStreamSynchronize(c->st));
                for (uint32_t i = 0; i < n2; i++) pass.push_back({base + c->h_k3[i], c->h_a3[i]});
            }
        }
        nph += n1;
        base += nb;
    }
    std::sort(pass.begin(), pass.end());
    if (n_phase) *n_phase = nph;
    *n_out = pass.size();
    const uint64_t nw = std::min<uint64_t>(cap, pass.size());
    for (uint64_t i = 0; i < nw; i++) {
        if (seeds) seeds[i] = dsx::mix64(start + pass[i].first);
        if (est) est[i] = pass[i].second;
        if (kidx) kidx[i] = pass[i].first;
    }
    return pass.size() > cap ? DSX_E_CAPACITY : DSX_OK;
}

int dsx_gate_amps(dsx_ctx *c, const uint64_t *seeds, uint64_t n, const dsx_gate_params *p, float *pre, float *full) {
    if (!c || !valid_gate_params(p) || (n && !seeds)) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    const dsx::GateKernel kpre = gate_kernel(4, p->box ? p->sig : 1.0f);
    const dsx::GateKernel kfull = gate_kernel(7, 2.0f);
    const int mode_pre = p->box ? 0 : 2;
    // (tests) DSX_GATE_AMPS_FAST = 1, 2 or 3: pre gets the fast pre-gate amplitudes of that decision kernel
    // (GATE_SPLIT, GATE_SPLIT_T, GATE_SPLIT_Q; the last two for box = 1 only) and full their margins;
    // DSX_GATE_AMPS_FAST_FULL = 1: full gets the full gate's fast amplitudes (GATE_SPLIT_Q) and pre their margins
    int fast_dec = dsx::GATE_EXACT;
    if (const char *e = std::getenv("DSX_GATE_AMPS_FAST"); e && *e) {
        fast_dec = std::atoi(e);
        if (fast_dec < dsx::GATE_SPLIT || fast_dec > dsx::GATE_SPLIT_Q || (fast_dec != dsx::GATE_SPLIT && !p->box))
            return DSX_E_ARG;
    }
    const char *ef = std::getenv("DSX_GATE_AMPS_FAST_FULL");
    const bool fast_full = ef && *ef == '1';
    if (fast_full) {
        if (fast_dec != dsx::GATE_EXACT || !pre || !full) return DSX_E_ARG;
        for (uint64_t b = 0; b < n; b += c->cap1) {
            const uint32_t m = uint32_t(std::min<uint64_t>(c->cap1, n - b));
            HIPCK(hipMemcpyAsync(c->d_seeds, seeds + b, size_t(m) * 8, hipMemcpyHostToDevice, c->st));
            HIPCK(hipMemsetAsync(c->d_cnt, 0, 32, c->st));
            dsx::launch_init(c->st, 0, nullptr, c->d_seeds, m, c->d_sides);
            dsx::launch_amp(c->st, 1, dsx::GATE_SPLIT_Q, c->d_sides, nullptr, m, kfull, 1e300, nullptr, c->d_i2, c->d_a3,
                            c->d_cnt, c->cap1, c->d_all, nullptr, c->d_k3, c->d_cnt + 2);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(full + b, c->d_all, size_t(m) * 4, hipMemcpyDeviceToHost, c->st));
            HIPCK(hipMemcpyAsync(pre + b, c->d_a3, size_t(m) * 4, hipMemcpyDeviceToHost, c->st));
            HIPCK(hipStreamSynchronize(c->st));
        }
        return DSX_OK;
    }
    for (uint64_t b = 0; b < n; b += c->cap1) {
        const uint32_t m = uint32_t(std::min<uint64_t>(c->cap1, n - b));
        HIPCK(hipMemcpyAsync(c->d_seeds, seeds + b, size_t(m) * 8, hipMemcpyHostToDevice, c->st));
        HIPCK(hipMemsetAsync(c->d_cnt, 0, 16, c->st));
        dsx::launch_init(c->st, 0, nullptr, c->d_seeds, m, c->d_sides);
        HIPCK(hipGetLastError());
        if (pre) {
            dsx::launch_amp(c->st, mode_pre, fast_dec, c->d_sides, nullptr, m, kpre, 1e300, nullptr, c->d_i2,
                            fast_dec ? c->d_a3 : nullptr, c->d_cnt, c->cap1, c->d_all, nullptr, c->d_k3, c->d_cnt + 2);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(pre + b, c->d_all, size_t(m) * 4, hipMemcpyDeviceToHost, c->st));
            if (fast_dec && full) HIPCK(hipMemcpyAsync(full + b, c->d_a3, size_t(m) * 4, hipMemcpyDeviceToHost, c->st));
        }
        if (full && !fast_dec) {
            dsx::launch_amp(c->st, 1, dsx::GATE_EXACT, c->d_sides, nullptr, m, kfull, 1e300, nullptr, c->d_i2, nullptr, c->d_cnt + 1,
                            c->cap1, c->d_a3);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(full + b, c->d_a3, size_t(m) * 4, hipMemcpyDeviceToHost, c->st));
        }
        HIPCK(hipStreamSynchronize(c->st));
    }
    return DSX_OK;
}

int dsx_gate_profile(dsx_ctx *c, double ms[4], uint64_t work[4]) {
    if (!c || !ms || !work) return DSX_E_ARG;
    for (int i = 0; i < 4; i++) { ms[i] = c->prof_ms[i]; work[i] = c->prof_work[i]; }
    return DSX_OK;
}

}  // extern "C"
