// Period tables and translation scan: shared declarations (tables.hip, scan host code).
#pragma once
#include "dsx_internal.h"

namespace dsx {

constexpr int TAB_BAND = 64;          // output rows per k_tab wave (8 bands per 512-row table)
constexpr int NT_TRANS = 60;          // world translations per axis (dsearch_desert.cpp:46)
constexpr int SCAN_KMIN = -30;        // first translation (dsearch_desert.cpp:47)
constexpr int SCAN_CELL = 512;        // quarts per T-table cell (dsearch_desert.cpp:48)
constexpr int SCAN_PERIOD = 262144;   // side-A period of temperature octave 0 in quarts (dsearch_desert.cpp:45)
constexpr int SCAN_WORLD = 7499996;   // world border in quarts (dsearch_desert.cpp:44)

struct TabArgs {
    float kT[17], kC[17];   // normalised blur weights float(exp(-i^2/2s^2) / sum), 2r+1 used
    double w[3];            // octave-0 weight factor * weight of T, C, E
};
struct ScanArgs {
    float ta, tt, wc, we;
};
struct ScanCand {
    float s, t, c, e;
    int x, z;
    uint32_t pos, pad_;
};

// zero[0..nzero): counters of the chunk's later kernels, cleared first
void launch_octinit(hipStream_t st, const uint64_t *seeds, uint32_t n, SideParams *out, uint32_t *zero = nullptr,
                    uint32_t nzero = 0);
bool launch_tab(hipStream_t st, const SideParams *sp, uint32_t n, int t0, int nt, int R, const TabArgs &A, float *tabs,
                float *bandmax);
// active cells of seed s to act[s * seg ..], counted in count[s]
void launch_active(hipStream_t st, const float *tabs, uint32_t n, float ta, uint32_t *act, uint32_t *count, uint32_t seg);
// the active cells of the m seeds (XCD x scans seeds x, x + 8, ...)
void launch_scan(hipStream_t st, const float *tabs, const uint32_t *act, const uint32_t *count, uint32_t m, uint32_t seg,
                 const int16_t *ibT, const int8_t *loT, const int8_t *hiT, const int16_t *ibC, const ScanArgs &A,
                 ScanCand *cand, uint32_t *ncand, uint32_t cap);
// per seed k of m: h_max[2k..2k+1] = max TA, max TB; h_n[1 + k] = candidates, h_n[1 + m + k] = active
// cells; up to pcap candidates to h_cand[k * pcap ..]
void launch_scan_publish(hipStream_t st, uint32_t m, const float *bandmax, const uint32_t *count, const uint32_t *ncand,
                         const ScanCand *cand, uint32_t cap, uint32_t pcap, float *h_max, uint32_t *h_n, ScanCand *h_cand);

}  // namespace dsx
