// Period tables, Gaussian blur and the world translation scan (stages S1-S2) for gfx950.
//
// Reference semantics:
//   octave-0 Perlin of T, C, E per side           standalone.hpp:123-139 (NormalNoise::init)
//   6 period tables (TA,TB,CA,CB,EA,EB), 512^2    lowfield.hpp:14-55 (tab_setup, tabulate_octave)
//   periodic separable blur, normalised weights   lowfield.hpp:105-139 (Blur)
//   maxta / maxtb                                 dsearch_desert.cpp:247
//   active side-A cells (TA >= ta) and the scan   dsearch_desert.cpp:256-280 (gproxy.cu:180-209)
// Each table is tabulated and blurred by one wave per row band (grid_blur.h, MODE 1: the period table
// is exactly the full gate's 512^2 grid with the channel's own octave weight), written once to HBM.
// The scan runs one wave per active side-A cell with lanes on the 60 x translations, so the side-B
// maps of a lane are loop invariant and each tz step is one gather per table; every candidate carries
// its scan position ((j*512 + i)*64 + tz)*64 + tx, from which the host restores the CPU's order.
#include "grid_blur.h"
#include "tables.h"

namespace dsx {

// Octave 0 of channel T/C/E (table t: channel t/2, side t&1) for every seed: one lane per table,
// permutation shuffled in a lane-private LDS slot (same scheme as k_init in gate.hip).
constexpr int PSTRIDE = 260;
__global__ __launch_bounds__(64) void k_octinit(const uint64_t *seeds, uint32_t n, SideParams *out, uint32_t *zero,
                                                 uint32_t nzero) {
    __shared__ uint32_t slot32[64 * PSTRIDE / 4];
    uint8_t *slots = reinterpret_cast<uint8_t *>(slot32);
    const int lane = threadIdx.x;
    const uint32_t g = blockIdx.x * 64u + lane;
    if (blockIdx.x == 0 && zero)   // counters the later kernels of the chunk accumulate into
        for (uint32_t k = lane; k < nzero; k += 64) zero[k] = 0;
    const bool active = g < 6u * n;
    uint8_t *my = slots + lane * PSTRIDE;
    double ox = 0, oy = 0, oz = 0;
    if (active) {
        const uint32_t s = g / 6, t = g % 6;
        const int ch = (t / 2 == 0) ? DSX_CH_T : (t / 2 == 1) ? DSX_CH_C : DSX_CH_E;
        Xoro w = Xoro::seeded(seeds[s]);
        const uint64_t a = w.next(), b = w.next();
        Xoro c = Xoro::raw(a ^ DSX_NOISE[ch].fork_a, b ^ DSX_NOISE[ch].fork_b);
        uint64_t ra = c.next(), rb = c.next();
        if (t & 1) { ra = c.next(); rb = c.next(); }
        const uint64_t *salt = DSX_OCT_SALT[DSX_NOISE[ch].first + 10];
        Xoro r = Xoro::raw(ra ^ salt[0], rb ^ salt[1]);
        ox = r.uniform() * 256;
        oy = r.uniform() * 256;
        oz = r.uniform() * 256;
        uint32_t *my32 = reinterpret_cast<uint32_t *>(my);
        for (int q = 0; q < 64; q++) my32[q] = 0x03020100u + 0x04040404u * uint32_t(q);
        for (int i = 0; i < 256; i++) {
            const int j = i + int(r.bounded(uint32_t(256 - i)));
            const uint8_t tmp = my[i];
            my[i] = my[j];
            my[j] = tmp;
        }
    }
    __syncthreads();
    const uint32_t g0 = blockIdx.x * 64u;
    for (int q = 0; q < 64; q++) {
        if (g0 + q >= 6u * n) break;
        reinterpret_cast<uint32_t *>(out[g0 + q].perm)[lane] = slot32[q * (PSTRIDE / 4) + lane];
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

// One wave per (seed, table, band of TAB_BAND output rows). first_table/ntab select T (0..1) or C/E
// (2..5) so that T and C/E may use different blur radii. bandmax gets the band's maximum.
template <int R>
__global__ __launch_bounds__(64) void k_tab(const SideParams *sp, uint32_t n, int first_table, int ntab, TabArgs A,
                                            float *tabs, float *bandmax) {
    __shared__ WaveLds L;
    constexpr int NB = 512 / TAB_BAND;
    const uint32_t blk = blockIdx.x;
    const uint32_t band = blk % NB, tq = (blk / NB) % ntab, s = blk / (NB * ntab);
    if (s >= n) return;
    const int t = first_table + int(tq);
    const SideParams &p = sp[size_t(s) * 6 + t];
    grid_setup<1>(p, L, A.w[t / 2]);
    const int lane = threadIdx.x;
    float *dst = tabs + (size_t(s) * 6 + t) * (512 * 512);
    float best = -1e30f;
    const float *kern = (t < 2) ? A.kT : A.kC;
    grid_blur_rows<1, R>(p, L, kern, int(band) * TAB_BAND, int(band) * TAB_BAND + TAB_BAND - 1,
                         [&](int o, const float (&v)[8]) {
                             float4 *row = reinterpret_cast<float4 *>(dst + size_t(o) * 512 + 8 * lane);
                             row[0] = make_float4(v[0], v[1], v[2], v[3]);
                             row[1] = make_float4(v[4], v[5], v[6], v[7]);
#pragma unroll
                             for (int q = 0; q < 8; q++) best = fmaxf(best, v[q]);
                         });
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o));
    if (lane == 0) bandmax[(size_t(s) * 6 + t) * NB + band] = best;
}

// Active cells (TA >= ta) of seed s = blockIdx.y into its own segment act[s * seg ..], counted in count[s]
// (cells past the segment are counted, not stored), in any order. A workgroup takes 2048 cells (8 per
// thread, coalesced) and reserves their slots with one atomic: one per wave with an active cell kept the
// kernel far above its 60 MB of reads, the atomics on each seed's counter serialising.
constexpr uint32_t ACT_PER_THREAD = 8, ACT_PER_WG = 256 * ACT_PER_THREAD;
__global__ __launch_bounds__(256) void k_active(const float *tabs, uint32_t n, float ta, uint32_t *act, uint32_t *count,
                                                uint32_t seg) {
    __shared__ uint32_t s_wtot[4], s_base;
    const uint32_t s = blockIdx.y;
    if (s >= n) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t c0 = blockIdx.x * ACT_PER_WG + threadIdx.x;
    const float *ta_tab = tabs + size_t(s) * 6 * 512 * 512;
    uint64_t bal[ACT_PER_THREAD];
    uint32_t wtot = 0;
#pragma unroll
    for (uint32_t i = 0; i < ACT_PER_THREAD; i++) {
        bal[i] = __ballot(ta_tab[c0 + i * 256u] >= ta);
        wtot += uint32_t(__popcll(bal[i]));
    }
    if (lane == 0) s_wtot[wave] = wtot;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = s_wtot[0] + s_wtot[1] + s_wtot[2] + s_wtot[3];
        s_base = tot ? atomicAdd(&count[s], tot) : 0u;
    }
    __syncthreads();
    if (wtot == 0) return;
    uint32_t slot = s_base;
    for (int w = 0; w < wave; w++) slot += s_wtot[w];
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (uint32_t i = 0; i < ACT_PER_THREAD; i++) {
        if ((bal[i] >> lane) & 1ull) {
            const uint32_t k = slot + uint32_t(__popcll(bal[i] & lt));
            if (k < seg) act[size_t(s) * seg + k] = c0 + i * 256u;
        }
        slot += uint32_t(__popcll(bal[i]));
    }
}

__device__ __forceinline__ void scan_cell(const float *tabs, uint32_t s, uint32_t cell, const int16_t *ibT,
                                          const int8_t *loT, const int8_t *hiT, const int16_t *ibC, const ScanArgs &A,
                                          ScanCand *cand, uint32_t *ncand, uint32_t cap, ScanCand *stage, uint32_t &nst);
// A wave's candidates are staged in LDS (SCAN_STAGE of them) and appended to seed s's list with one atomic
// per flush: one atomic per translation step with a hit, on the counter of the one seed an XCD scans at a
// time, serialised. The list is in any order either way (the host sorts it by scan position).
constexpr uint32_t SCAN_STAGE = 64;
__device__ __forceinline__ void scan_flush(uint32_t s, ScanCand *stage, uint32_t &nst, ScanCand *cand, uint32_t *ncand,
                                           uint32_t cap) {
    if (nst == 0) return;
    const uint32_t lane = threadIdx.x;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&ncand[s], nst);
    base = __shfl(base, 0);
    __syncthreads();   // (one wave per workgroup) the staged entries are written
    if (lane < nst && base + lane < cap) cand[size_t(s) * cap + base + lane] = stage[lane];
    __syncthreads();   // and read before the stage is reused
    nst = 0;
}
// XCD-aware: workgroups are dispatched to the 8 XCDs round-robin (XCD = blockIdx.x % 8), and the
// workgroups of XCD x scan seeds x, x + 8, ... one after another, striding over each seed's active
// cells, so each XCD's L2 holds the B-side tables (TB, CB, EB: 3 MB) of one seed at a time instead of
// the gathers of all seeds of the chunk missing to memory.
__global__ __launch_bounds__(64) void k_scan(const float *tabs, const uint32_t *act, const uint32_t *count, uint32_t m,
                                             uint32_t seg, const int16_t *ibT, const int8_t *loT, const int8_t *hiT,
                                             const int16_t *ibC, ScanArgs A, ScanCand *cand, uint32_t *ncand,
                                             uint32_t cap) {
    __shared__ ScanCand stage[SCAN_STAGE];
    uint32_t nst = 0;
    constexpr uint32_t NXCD = 8;
    const uint32_t xcd = blockIdx.x % NXCD, lb = blockIdx.x / NXCD, per = gridDim.x / NXCD;
    for (uint32_t s = xcd; s < m; s += NXCD) {
        const uint32_t na = min(count[s], seg);
        for (uint32_t a = lb; a < na; a += per)
            scan_cell(tabs, s, act[size_t(s) * seg + a], ibT, loT, hiT, ibC, A, cand, ncand, cap, stage, nst);
        scan_flush(s, stage, nst, cand, ncand, cap);
    }
}

// Per seed of the chunk (one block each): max of the blurred TA / TB tables from their band maxima,
// the candidate count, and up to pcap candidates, into host memory; out_n = the active-cell count.
__global__ __launch_bounds__(64) void k_scan_publish(const float *bandmax, const uint32_t *count, const uint32_t *ncand,
                                                     const ScanCand *cand, uint32_t cap, uint32_t pcap, float *h_max,
                                                     uint32_t *h_n, ScanCand *h_cand) {
    // h_n[1 + k] candidates of seed k; h_n[1 + m + k] its active cells (m = gridDim.x); h_n[0] unused
    const uint32_t k = blockIdx.x;
    const int lane = threadIdx.x;
    constexpr int NB = 512 / TAB_BAND;
    if (lane < 2) {
        float m = -1e30f;
        for (int q = 0; q < NB; q++) m = fmaxf(m, bandmax[(size_t(k) * 6 + lane) * NB + q]);
        h_max[2 * k + lane] = m;
    }
    const uint32_t nc = ncand[k];
    if (lane == 0) {
        h_n[1 + k] = nc;
        h_n[1 + gridDim.x + k] = count[k];
    }
    const uint32_t w = min(min(nc, cap), pcap);
    constexpr uint32_t W = sizeof(ScanCand) / 4;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(cand + size_t(k) * cap);
    uint32_t *dst = reinterpret_cast<uint32_t *>(h_cand + size_t(k) * pcap);
    for (uint32_t q = lane; q < w * W; q += 64) dst[q] = src[q];
}

__device__ __forceinline__ void scan_cell(const float *tabs, uint32_t s, uint32_t cell, const int16_t *ibT,
                                          const int8_t *loT, const int8_t *hiT, const int16_t *ibC, const ScanArgs &A,
                                          ScanCand *cand, uint32_t *ncand, uint32_t cap, ScanCand *stage, uint32_t &nst) {
    const int j = int(cell >> 9), i = int(cell & 511), lane = threadIdx.x;
    const float *TA = tabs + size_t(s) * 6 * 512 * 512;
    const float *TB = TA + 512 * 512, *CA = TB + 512 * 512, *CB = CA + 512 * 512, *EA = CB + 512 * 512,
                *EB = EA + 512 * 512;
    const float sa = TA[cell];
    const int ca = ((2 * j) & 511) * 512 + ((2 * i) & 511);
    const float cA = CA[ca], eA = EA[ca];
    const bool lane_ok = lane < NT_TRANS && lane >= loT[i] && lane <= hiT[i];
    const int cT = lane_ok ? ibT[i * NT_TRANS + lane] : 0, cC = lane_ok ? ibC[i * NT_TRANS + lane] : 0;
    const float tt = A.tt, pre = A.tt - 0.3f, wc = A.wc, we = A.we;
    const int tz0 = loT[j], tz1 = hiT[j];
    for (int tz = tz0; tz <= tz1; tz++) {
        const int r = ibT[j * NT_TRANS + tz], rc = ibC[j * NT_TRANS + tz];
        bool hit = false;
        float tv = 0, cv = 0, ev = 0, sc = 0;
        if (lane_ok) {
            tv = sa + TB[r * 512 + cT];
            if (!(tv < pre)) {
                cv = cA + CB[rc * 512 + cC];
                ev = eA + EB[rc * 512 + cC];
                sc = tv + wc * cv + we * ev;
                hit = !(sc < tt);
            }
        }
        const uint64_t mask = __ballot(hit);
        if (mask == 0) continue;
        const uint32_t nh = uint32_t(__popcll(mask));
        if (nst + nh > SCAN_STAGE) scan_flush(s, stage, nst, cand, ncand, cap);
        if (hit) {
            const uint32_t k = nst + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
            {
                ScanCand &o = stage[k];
                o.s = sc;
                o.t = tv;
                o.c = cv;
                o.e = ev;
                o.x = i * SCAN_CELL + (lane + SCAN_KMIN) * SCAN_PERIOD;
                o.z = j * SCAN_CELL + (tz + SCAN_KMIN) * SCAN_PERIOD;
                o.pos = ((uint32_t(j) * 512u + uint32_t(i)) * 64u + uint32_t(tz)) * 64u + uint32_t(lane);
            }
        }
        nst += nh;
    }
}

// ------------------------------------------------------------------ launchers
void launch_octinit(hipStream_t st, const uint64_t *seeds, uint32_t n, SideParams *out, uint32_t *zero, uint32_t nzero) {
    if (n) hipLaunchKernelGGL(k_octinit, dim3((6 * n + 63) / 64), dim3(64), 0, st, seeds, n, out, zero, nzero);
}
template <int R>
static void tab_r(hipStream_t st, const SideParams *sp, uint32_t n, int t0, int nt, const TabArgs &A, float *tabs,
                  float *bandmax) {
    hipLaunchKernelGGL(k_tab<R>, dim3(n * nt * (512 / TAB_BAND)), dim3(64), 0, st, sp, n, t0, nt, A, tabs, bandmax);
}
bool launch_tab(hipStream_t st, const SideParams *sp, uint32_t n, int t0, int nt, int R, const TabArgs &A, float *tabs,
                float *bandmax) {
    if (!n) return true;
    switch (R) {
    case 4: tab_r<4>(st, sp, n, t0, nt, A, tabs, bandmax); return true;
    case 5: tab_r<5>(st, sp, n, t0, nt, A, tabs, bandmax); return true;
    case 6: tab_r<6>(st, sp, n, t0, nt, A, tabs, bandmax); return true;
    case 7: tab_r<7>(st, sp, n, t0, nt, A, tabs, bandmax); return true;
    case 8: tab_r<8>(st, sp, n, t0, nt, A, tabs, bandmax); return true;
    default: return false;
    }
}
void launch_active(hipStream_t st, const float *tabs, uint32_t n, float ta, uint32_t *act, uint32_t *count, uint32_t seg) {
    if (n) hipLaunchKernelGGL(k_active, dim3(512 * 512 / ACT_PER_WG, n), dim3(256), 0, st, tabs, n, ta, act, count, seg);
}
void launch_scan(hipStream_t st, const float *tabs, const uint32_t *act, const uint32_t *count, uint32_t m, uint32_t seg,
                 const int16_t *ibT, const int8_t *loT, const int8_t *hiT, const int16_t *ibC, const ScanArgs &A,
                 ScanCand *cand, uint32_t *ncand, uint32_t cap) {
    if (m) hipLaunchKernelGGL(k_scan, dim3(8 * 256), dim3(64), 0, st, tabs, act, count, m, seg, ibT, loT, hiT, ibC, A, cand,
                              ncand, cap);
}
void launch_scan_publish(hipStream_t st, uint32_t m, const float *bandmax, const uint32_t *count, const uint32_t *ncand,
                         const ScanCand *cand, uint32_t cap, uint32_t pcap, float *h_max, uint32_t *h_n, ScanCand *h_cand) {
    if (m)
        hipLaunchKernelGGL(k_scan_publish, dim3(m), dim3(64), 0, st, bandmax, count, ncand, cand, cap, pcap, h_max, h_n,
                           h_cand);
}

}  // namespace dsx
