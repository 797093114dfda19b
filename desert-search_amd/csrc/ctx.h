// Host-side context of libdsx.so, shared by the API translation units.
#pragma once
#include "../../include/dsx.h"
#include "dsx_internal.h"
#include "tables.h"
#include <cstdio>
#include <string>

namespace dsx {
extern thread_local std::string g_last_error;
int hip_fail(hipError_t e, const char *what, const char *file, int line);
GateKernel gate_kernel(int R, float sig);            // GPU-gate blur weights (gate.cuh:216, 248)
bool valid_gate_params(const dsx_gate_params *p);
}  // namespace dsx

#define HIPCK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) return dsx::hip_fail(e_, #x, __FILE__, __LINE__);      \
    } while (0)

void dsx_ctx_free_ext(struct dsx_ctx *c);      // climate / flood state (floods_api.hip)
void dsx_ctx_free_workers(struct dsx_ctx *c);  // pipeline worker contexts (host/search.cpp)
int dsx_ctx_device(const struct dsx_ctx *c);

struct dsx_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    // seed gate, sized for one batch
    uint32_t cap1 = 0;
    uint32_t *d_k1 = nullptr, *d_i2 = nullptr, *d_k3 = nullptr, *d_cnt = nullptr;
    float *d_a3 = nullptr, *d_all = nullptr;
    float *d_band = nullptr;   // gate graph: band maxima of undecided pairs (cap1 * 16), allocated with the graph
    dsx::SideParams *d_sides = nullptr;
    uint64_t *d_seeds = nullptr;
    uint32_t *h_cnt = nullptr, *h_k3 = nullptr;
    float *h_a3 = nullptr;
    hipEvent_t ev[8] = {};
    double prof_ms[4] = {0, 0, 0, 0};
    uint64_t prof_work[4] = {0, 0, 0, 0};
    // the gate as a replayed hipGraph (dsx_set_gate_graph): its configuration and the pinned outputs its
    // sort kernel writes (4 counts, then cap1 indices and cap1 amplitudes)
    int gate_graph = -1;                 // -1: from DSX_GATE_GRAPH (default on)
    dsx::GateGraph ggraph;
    uint64_t *h_gout = nullptr;
    // period tables + scan (allocated on first use)
    uint32_t scan_seeds = 0;             // seeds per chunk
    dsx::SideParams *d_tsp = nullptr;    // 6 octave sides per seed
    float *d_tabs = nullptr;             // 6 x 512 x 512 floats per seed
    float *d_bandmax = nullptr;
    uint32_t *d_act = nullptr, *d_ncand = nullptr;
    uint32_t act_cap = 0;
    dsx::ScanCand *d_cand = nullptr;
    uint64_t cand_cap = 0;               // candidates per seed the buffer holds
    int16_t *d_ibT = nullptr, *d_ibC = nullptr;
    int8_t *d_loT = nullptr, *d_hiT = nullptr;
    // pinned scan I/O: seeds up; per-seed maxima, counts and candidates down (one synchronisation)
    uint64_t *h_seeds = nullptr;
    float *h_smax = nullptr;
    uint32_t *h_sn = nullptr;
    dsx::ScanCand *h_cand = nullptr;
};
