// C ABI for stages S1-S2: period tables, blur, maxta/maxtb and the raw translation-scan candidates
// (include/dsx.h dsx_scan, dsx_tables). Replaces gproxy::scan_init / gproxy::scan (gproxy.h:30-33).
#include "ctx.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr uint32_t SCAN_CHUNK = 64;   // seeds per table batch (6 MB of tables each)

// Translation map of an octave whose side-A lattice unit is `lat` quarts: for side-A cell i and
// translation t, the side-B table cell, and the translations keeping the cell 6000 quarts inside the
// world (dsearch_desert.cpp:98-113).
void trans_map(double lat, int16_t *ib, int8_t *lo, int8_t *hi) {
    const double bs = 337.0 / 331.0;
    for (int i = 0; i < 512; i++) {
        int l = dsx::NT_TRANS, h = -1;
        for (int t = 0; t < dsx::NT_TRANS; t++) {
            const long x = long(i) * dsx::SCAN_CELL + long(t + dsx::SCAN_KMIN) * dsx::SCAN_PERIOD;
            double u = x * bs / lat;
            u -= std::floor(u / 256.0) * 256.0;
            ib[i * dsx::NT_TRANS + t] = int16_t(int(u * (512 / 256.0) + 0.5) & 511);
            if (x >= -dsx::SCAN_WORLD + 6000 && x <= dsx::SCAN_WORLD - 6000) {
                l = std::min(l, t);
                h = std::max(h, t);
            }
        }
        lo[i] = int8_t(l);
        hi[i] = int8_t(h);
    }
}

// Normalised Gaussian blur of radius int(3 sigma + 1) (lowfield.hpp:105-112).
int blur_weights(double sigma, float *k) {
    const int r = int(3 * sigma + 1);
    if (r < 4 || r > 8) return -1;
    double kd[17], s = 0;
    for (int i = -r; i <= r; i++) {
        kd[i + r] = std::exp(-0.5 * i * i / (sigma * sigma));
        s += kd[i + r];
    }
    for (int i = 0; i <= 2 * r; i++) k[i] = float(kd[i] / s);
    return r;
}

// Octave-0 weight (factor * weight) of channel ch (standalone.hpp:123-139).
double octave0_weight(int ch) {
    const dsx_noise_cfg &c = DSX_NOISE[ch];
    const double w0 = double(1 << (c.len - 1)) / ((1 << c.len) - 1);
    int lo = 0, hi = c.len - 1;
    while (c.amp[lo] == 0) lo++;
    while (c.amp[hi] == 0) hi--;
    const int span = hi - lo + 1;
    const double factor = (5. * span) / (3. * (span + 1));
    return factor * (w0 * c.amp[0]);
}

int scan_alloc(dsx_ctx *c) {
    if (c->d_tabs) return DSX_OK;
    c->scan_seeds = SCAN_CHUNK;
    const size_t S = c->scan_seeds;
    HIPCK(hipMalloc(&c->d_tsp, S * 6 * sizeof(dsx::SideParams)));
    HIPCK(hipMalloc(&c->d_tabs, S * 6 * 512 * 512 * sizeof(float)));
    HIPCK(hipMalloc(&c->d_bandmax, S * 6 * (512 / dsx::TAB_BAND) * sizeof(float)));
    c->act_cap = 1u << 21;
    HIPCK(hipMalloc(&c->d_act, size_t(c->act_cap) * 8));
    HIPCK(hipMalloc(&c->d_ncand, 2 * S * 4));   // per-seed candidate counts, then per-seed active-cell counts
    HIPCK(hipHostMalloc(&c->h_seeds, S * 8, hipHostMallocDefault));
    HIPCK(hipHostMalloc(&c->h_smax, S * 2 * 4, hipHostMallocDefault));
    HIPCK(hipHostMalloc(&c->h_sn, (2 * S + 1) * 4, hipHostMallocDefault));
    std::vector<int16_t> ibT(512 * dsx::NT_TRANS), ibC(512 * dsx::NT_TRANS), ibX(512 * dsx::NT_TRANS);
    std::vector<int8_t> loT(512), hiT(512), loC(512), hiC(512);
    trans_map(1024.0, ibT.data(), loT.data(), hiT.data());
    trans_map(512.0, ibC.data(), loC.data(), hiC.data());
    HIPCK(hipMalloc(&c->d_ibT, ibT.size() * 2));
    HIPCK(hipMalloc(&c->d_ibC, ibC.size() * 2));
    HIPCK(hipMalloc(&c->d_loT, 512));
    HIPCK(hipMalloc(&c->d_hiT, 512));
    // ordered on the context's stream (a plain hipMemcpy goes to the null stream) and completed before
    // the host buffers go out of scope
    HIPCK(hipMemcpyAsync(c->d_ibT, ibT.data(), ibT.size() * 2, hipMemcpyHostToDevice, c->st));
    HIPCK(hipMemcpyAsync(c->d_ibC, ibC.data(), ibC.size() * 2, hipMemcpyHostToDevice, c->st));
    HIPCK(hipMemcpyAsync(c->d_loT, loT.data(), 512, hipMemcpyHostToDevice, c->st));
    HIPCK(hipMemcpyAsync(c->d_hiT, hiT.data(), 512, hipMemcpyHostToDevice, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    return DSX_OK;
}

int cand_alloc(dsx_ctx *c, uint64_t per_seed) {
    if (c->d_cand && c->cand_cap >= per_seed) return DSX_OK;
    if (c->d_cand) HIPCK(hipFree(c->d_cand));
    c->d_cand = nullptr;
    c->cand_cap = std::max<uint64_t>({per_seed, 2 * c->cand_cap, 1u << 12});   // doubles: hipFree waits for the device
    HIPCK(hipMalloc(&c->d_cand, size_t(SCAN_CHUNK) * c->cand_cap * sizeof(dsx::ScanCand)));
    if (c->h_cand) HIPCK(hipHostFree(c->h_cand));
    c->h_cand = nullptr;
    HIPCK(hipHostMalloc(&c->h_cand, size_t(SCAN_CHUNK) * c->cand_cap * sizeof(dsx::ScanCand), hipHostMallocDefault));
    return DSX_OK;
}

bool valid_scan_params(const dsx_scan_params *p) {
    return p && std::isfinite(p->ta) && std::isfinite(p->tt) && std::isfinite(p->wc) && std::isfinite(p->we) &&
           std::isfinite(p->sigma) && std::isfinite(p->sigmac);
}

// Tables of seeds[0..m) into c->d_tabs; band maxima in c->d_bandmax.
int tables_chunk(dsx_ctx *c, const uint64_t *seeds, uint32_t m, const dsx_scan_params *p) {
    dsx::TabArgs A{};
    const int rT = blur_weights(p->sigma, A.kT), rC = blur_weights(p->sigmac, A.kC);
    if (rT < 0 || rC < 0) {
        dsx::g_last_error = "blur radius int(3*sigma+1) must be in [4, 8]";
        return DSX_E_ARG;
    }
    A.w[0] = octave0_weight(DSX_CH_T);
    A.w[1] = octave0_weight(DSX_CH_C);
    A.w[2] = octave0_weight(DSX_CH_E);
    // the init kernel reads the seeds from pinned memory in place (callers synchronise the stream before
    // the next chunk overwrites them)
    std::memcpy(c->h_seeds, seeds, size_t(m) * 8);
    dsx::launch_octinit(c->st, c->h_seeds, m, c->d_tsp, c->d_ncand, 2 * m);
    HIPCK(hipGetLastError());
    dsx::launch_tab(c->st, c->d_tsp, m, 0, 2, rT, A, c->d_tabs, c->d_bandmax);
    HIPCK(hipGetLastError());
    dsx::launch_tab(c->st, c->d_tsp, m, 2, 4, rC, A, c->d_tabs, c->d_bandmax);
    HIPCK(hipGetLastError());
    return DSX_OK;
}

}  // namespace

extern "C" {

int dsx_tables(dsx_ctx *c, uint64_t seed, const dsx_scan_params *p, float *out) {
    if (!c || !valid_scan_params(p) || !out) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    int rc = scan_alloc(c);
    if (rc) return rc;
    rc = tables_chunk(c, &seed, 1, p);
    if (rc) return rc;
    HIPCK(hipMemcpyAsync(out, c->d_tabs, size_t(6) * 512 * 512 * 4, hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    return DSX_OK;
}

int dsx_scan(dsx_ctx *c, const uint64_t *seeds, uint64_t n, const dsx_scan_params *p, float *maxta, float *maxtb,
             dsx_raw_cand *cands, uint64_t cap, uint64_t *ncand) {
    if (!c || !valid_scan_params(p) || (n && (!seeds || !ncand)) || (cap && !cands)) return DSX_E_ARG;
    HIPCK(hipSetDevice(c->device));
    int rc = scan_alloc(c);
    if (rc) return rc;
    rc = cand_alloc(c, std::max<uint64_t>(cap, 1));
    if (rc) return rc;
    const dsx::ScanArgs A{p->ta, p->tt, p->wc, p->we};
    std::vector<dsx::ScanCand> buf;
    bool overflow = false;
    const uint32_t kcap = uint32_t(std::min<uint64_t>(c->cand_cap, 0xffffffffu));
    for (uint64_t b = 0; b < n;) {
        uint32_t m = uint32_t(std::min<uint64_t>(c->scan_seeds, n - b));
        rc = tables_chunk(c, seeds + b, m, p);   // also clears the chunk's counters
        if (rc) return rc;
        uint32_t *d_nact = c->d_ncand + m;
        const uint32_t seg = c->act_cap / m;   // active-cell segment of each seed
        dsx::launch_active(c->st, c->d_tabs, m, p->ta, c->d_act, d_nact, seg);
        dsx::launch_scan(c->st, c->d_tabs, c->d_act, d_nact, m, seg, c->d_ibT, c->d_loT, c->d_hiT, c->d_ibC, A, c->d_cand,
                         c->d_ncand, kcap);
        dsx::launch_scan_publish(c->st, m, c->d_bandmax, d_nact, c->d_ncand, c->d_cand, kcap, kcap, c->h_smax, c->h_sn,
                                 c->h_cand);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->st));
        uint32_t nact = 0;
        for (uint32_t k = 0; k < m; k++) nact = std::max(nact, c->h_sn[1 + m + k]);
        if (nact > seg) {                   // a seed's active cells overflow its segment: fewer seeds
            if (m == 1) {
                dsx::g_last_error = "active-cell list overflow for a single seed";
                return DSX_E_CAPACITY;
            }
            c->scan_seeds = std::max<uint32_t>(1, m / 2);
            continue;
        }
        for (uint32_t k = 0; k < m; k++) {
            if (maxta) maxta[b + k] = c->h_smax[2 * k];
            if (maxtb) maxtb[b + k] = c->h_smax[2 * k + 1];
            const uint32_t nc = c->h_sn[1 + k];
            ncand[b + k] = nc;
            const uint32_t have = std::min(nc, kcap);
            if (nc > cap) overflow = true;
            buf.assign(c->h_cand + size_t(k) * kcap, c->h_cand + size_t(k) * kcap + have);
            std::sort(buf.begin(), buf.end(), [](const dsx::ScanCand &x, const dsx::ScanCand &y) { return x.pos < y.pos; });
            const uint64_t w = std::min<uint64_t>(have, cap);
            dsx_raw_cand *dst = cands + (b + k) * cap;
            for (uint64_t q = 0; q < w; q++) dst[q] = {buf[q].s, buf[q].t, buf[q].c, buf[q].e, buf[q].x, buf[q].z};
        }
        b += m;
    }
    return overflow ? DSX_E_CAPACITY : DSX_OK;
}

}  // extern "C"
