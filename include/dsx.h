/*
 * dsx.h - C ABI of the MI355X desert-search engine (libdsx.so, gfx950 HIP kernels).
 *
 * This is the drop-in boundary for desert-search's per-seed pipeline. Each entry point replaces a
 * function of the reference's GPU engine, whose C++ interface lives in src/gpu/gproxy.h and
 * src/gpu/gate.cuh; the reference file:line is cited on each declaration. Unlike the reference:
 *   - every call takes an explicit per-device context (the reference keeps file-scope statics);
 *   - errors are integer codes (DSX_OK == 0), never exit(); dsx_strerror() names them;
 *   - outputs go to caller-owned buffers with explicit capacities; overflow is DSX_E_CAPACITY with the
 *     full count returned, never a silent truncation (gate.cuh:387,395,405 truncate silently);
 *   - results are canonical: survivor lists are sorted by seed index, so they do not depend on GPU
 *     scheduling (the reference's atomic appends do).
 * A context may be used by one host thread at a time; distinct contexts are independent.
 * No torch types, no STL: plain pointers and sizes only.
 */
#ifndef DSX_H
#define DSX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
enum {
    DSX_OK = 0,
    DSX_E_ARG = 1,       /* invalid argument (null pointer, bad size or mode) */
    DSX_E_HIP = 2,       /* a HIP runtime call failed (message via dsx_last_error) */
    DSX_E_CAPACITY = 3,  /* caller buffer or internal queue too small; *n_out holds the true count */
    DSX_E_NODEVICE = 4,  /* no usable gfx950 device */
    DSX_E_BUSY = 5,      /* the request has not finished yet (dsx_poll) */
    DSX_E_INTERRUPTED = 6, /* dsx_search stopped early after dsx_interrupt(1); its counts cover the seeds it wrote */
};
const char *dsx_strerror(int code);
/* Text of the last HIP failure on this thread (empty if none). */
const char *dsx_last_error(void);

/* ---------------------------------------------------------------- context */
typedef struct dsx_ctx dsx_ctx;
/* Creates the engine on HIP device `device` and uploads the constant tables.
 * Replaces gproxy::init (gproxy.h:35, gproxy.cu:391-401) minus its detached thread. */
int dsx_create(int device, dsx_ctx **out);
int dsx_destroy(dsx_ctx *ctx);

/* ---------------------------------------------------------------- seed gate (S0) */
typedef struct {
    float max_phase; /* y-phase cut, pass if phase <= (double)max_phase   (GateCfg::maxph, 0.06 in production) */
    float amp;       /* full 512^2 amplitude gate threshold               (GateCfg::amp, 0.70) */
    float pthr;      /* anti-aliased 256^2 pre-gate threshold, box=1      (GateCfg::pthr, 0.645) */
    float margin;    /* legacy pre-gate threshold amp - margin, box=0     (GateCfg::margin, 0.01) */
    float sig;       /* pre-gate blur sigma in grid cells                 (GateCfg::sig, 1.0) */
    int box;         /* 1: 4-parity anti-aliased pre-gate; 0: legacy even-parity pre-gate */
} dsx_gate_params;

typedef struct {
    uint64_t phase; /* seeds passing the y-phase filter */
    uint64_t pre;   /* ... and the pre-gate */
    uint64_t full;  /* ... and the full amplitude gate */
} dsx_gate_counts;

/* The GPU seed gate over seed indices k in [start, start+count), seed = splitmix64(start+k):
 * y-phase filter -> pre-gate -> full 512^2 amplitude gate.
 * Replaces tg::run_gate (gate.cuh:361-419) and gproxy::start_gate (gproxy.h:40-41).
 * Survivors are written sorted by k: seeds[i] (the seed), amps[i] (full amplitude A+B), kidx[i] (k).
 * Any of seeds/amps/kidx may be null. *n_out = number of survivors; DSX_E_CAPACITY if > cap. */
int dsx_gate(dsx_ctx *ctx, uint64_t start, uint64_t count, const dsx_gate_params *p, uint64_t *seeds, float *amps,
             uint64_t *kidx, uint64_t cap, uint64_t *n_out, dsx_gate_counts *counts);

/* The CPU-default gate of the reference's seed enumeration (dsearch_desert.cpp:202-223, gpu=0 or
 * gate=0) on the GPU: seed = splitmix64(start+k), keep k if phase <= max_phase (double compare,
 * dsearch_desert.cpp:206) and amplitude_sub A + B >= amp - margin (lowfield.hpp:60-102,
 * dsearch_desert.cpp:217-219). Survivors sorted by k: seeds[i], est[i] (the float estimate A + B),
 * kidx[i]; *n_phase = phase survivors (the `passed` counter). DSX_E_CAPACITY if more than cap. */
int dsx_gate_cpu(dsx_ctx *ctx, uint64_t start, uint64_t count, double max_phase, double amp, double margin,
                 uint64_t *seeds, float *est, uint64_t *kidx, uint64_t cap, uint64_t *n_out, uint64_t *n_phase);

/* Pre-gate and full-gate amplitudes (A+B) of an explicit seed list, no thresholds applied.
 * Replaces the calibration mode of the reference's gate tool (tgate.cu:53-85, `calib=N`). */
int dsx_gate_amps(dsx_ctx *ctx, const uint64_t *seeds, uint64_t n, const dsx_gate_params *p, float *pre, float *full);

/* The gate of dsx_gate without any allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It takes no context; every intermediate lives in a static device array of the library
 * (DSX_NOALLOC_SCRATCH_BYTES, about 40 MB, held on each device the call runs on until the process ends).
 * Seed indices [start, start+count) with count <= DSX_NOALLOC_MAX_COUNT (larger: DSX_E_ARG); the call
 * enqueues its kernels on `stream` (a hipStream_t of `device`; null: the null stream) and returns without
 * waiting. Outputs are DEVICE memory written by those kernels: survivors sorted by k into seeds / amps /
 * kidx (each may be null; at most cap entries), and counts[0..3] = {phase survivors, pre-gate survivors,
 * full-gate survivors, entries written = min(full, cap)}. If counts[0] > DSX_NOALLOC_MAX_PHASE the phase
 * survivors overflowed the scratch list: no survivors are written (counts[1..3] = 0); call again on a
 * smaller range. Results equal dsx_gate's on the same range. All calls on one device share its scratch
 * array: calls that overlap in time on one device (other streams, other threads) race. */
enum {
    DSX_NOALLOC_MAX_COUNT = 1 << 22,    /* seed indices per call */
    DSX_NOALLOC_MAX_PHASE = 1 << 16,    /* phase survivors per call (0.72% of 2^22 at max_phase 0.06 is ~30k) */
    DSX_NOALLOC_SCRATCH_BYTES = 39846144
};
int dsx_gate_noalloc(int device, void *stream, uint64_t start, uint64_t count, const dsx_gate_params *p, uint64_t *seeds,
                     float *amps, uint64_t *kidx, uint64_t cap, uint64_t *counts);

/* How dsx_gate runs (no counterpart in the reference, whose run_gate synchronises after each stage,
 * gate.cuh:384-407): on = 1 (the default; DSX_GATE_GRAPH=0 changes the default) replays, per chunk, a
 * hipGraph of the whole gate (the stages read their survivor counts on the device; one synchronisation per
 * chunk), built once per context and shape: chunks are the largest power of two up to 2^26 indices whose
 * expected phase survivors fit the survivor buffers (2^24 at max_phase 0.06), or the smallest holding the
 * call. The graph's pre-gate decides a seed from a faster amplitude (fused multiply-add blurs) when that
 * amplitude is farther from the threshold than its proven error bound, and from the reference's arithmetic
 * otherwise, so every decision is the reference's; the full gate and every amplitude returned use the
 * reference's arithmetic. on = 0 launches the stages one by one in the reference's arithmetic and reads each
 * stage's survivor count back first (dsx_gate_profile times each kernel only then). Results are identical. */
int dsx_set_gate_graph(dsx_ctx *ctx, int on);

/* Device time (ms) of the kernels of the last dsx_gate call, summed over its batches (with the graph off):
 * ms[0] phase filter, ms[1] Perlin init, ms[2] pre-gate, ms[3] full gate; work[0..3] the number of
 * items each processed (seed indices, sides, pre-gate seeds, full-gate seeds). */
int dsx_gate_profile(dsx_ctx *ctx, double ms[4], uint64_t work[4]);

/* ---------------------------------------------------------------- period tables + scan (S1-S2) */
typedef struct {
    float ta;      /* side-A T prune: scan cells with TA >= ta        (Params::ta, 0.25) */
    float tt;      /* candidate score threshold                        (Params::t, 0.55) */
    float wc, we;  /* score weights of smoothed C and E                (Params::wc 0.25, we 0.41) */
    double sigma;  /* T blur sigma in table cells                      (Params::sigma, 2.0) */
    double sigmac; /* C/E blur sigma in table cells                    (Params::sigmac, 2.0) */
} dsx_scan_params;

/* One raw candidate: score, its T/C/E parts and the world quart position (gproxy::RawCand, gproxy.h:28). */
typedef struct {
    float s, t, c, e;
    int32_t x, z;
} dsx_raw_cand;

/* Octave-0 period tables, blur and translation scan of each seed (seeds from the gate).
 * Replaces gproxy::scan_init + gproxy::scan (gproxy.h:30-33, gproxy.cu:221-307); the translation maps
 * and blur weights are derived internally from the parameters (dsearch_desert.cpp:98-113,
 * lowfield.hpp:105-112). Per seed i: maxta[i], maxtb[i] (max of the blurred side-A/side-B T tables),
 * ncand[i] = number of raw candidates (score >= tt), written to cands[i*cap .. i*cap + min(ncand, cap))
 * in the CPU scan order (j, i, tz, tx) of dsearch_desert.cpp:257-279, so the caller's bucket
 * reduction reproduces the CPU's candidate list. DSX_E_CAPACITY if any ncand[i] > cap. */
int dsx_scan(dsx_ctx *ctx, const uint64_t *seeds, uint64_t n, const dsx_scan_params *p, float *maxta, float *maxtb,
             dsx_raw_cand *cands, uint64_t cap, uint64_t *ncand);

/* The 6 blurred period tables {TA, TB, CA, CB, EA, EB} of one seed, 512*512 floats each, row-major
 * (lowfield.hpp tabulate_octave + Blur; the `gscancheck` comparison, dsearch_desert.cpp:292-296). */
int dsx_tables(dsx_ctx *ctx, uint64_t seed, const dsx_scan_params *p, float *out);

/* ---------------------------------------------------------------- FP64 climate sampler */
/* Noise channels (standalone.hpp:93) and flood predicates. */
enum { DSX_NOISE_T = 0, DSX_NOISE_H = 1, DSX_NOISE_C = 2, DSX_NOISE_E = 3, DSX_NOISE_SHIFT = 4, DSX_NOISE_W = 5 };
enum {
    DSX_PRED_DESERT_STATE = 0, /* SurfaceSampler::desert_state: 2 desert, 1 bridgeable river, 0 other (surface.hpp:98-124) */
    DSX_PRED_HOT = 1,          /* T(x, 0, z) >= 0.55 -> 2 else 0 (dsearch_desert.cpp:334) */
    DSX_PRED_PROXY = 2,        /* DesertProxy(cut): 2 or 0 (desert_proxy.hpp:32) */
    DSX_PRED_DESERT = 3        /* SurfaceSampler::desert -> 2 or 0 (surface.hpp:85-93) */
};

/* NormalNoise::sample (standalone.hpp:140-151) of `channel` for seed at n points xyz[3k..3k+2], FP64. */
int dsx_noise(dsx_ctx *ctx, uint64_t seed, int channel, const double *xyz, uint64_t n, double *out);
/* The same with the FP32 octaves of the pre-pass (lattice cells and fractions still FP64); no counterpart
 * in the reference: it measures the pre-pass error against dsx_noise. */
int dsx_noise_f32(dsx_ctx *ctx, uint64_t seed, int channel, const double *xyz, uint64_t n, double *out);
/* Precision of the context's flood and predicate kernels (no counterpart in the reference, which is FP64
 * throughout): 0 plain FP64; 1 FP32 pre-pass + FP64 confirm — the temperature test of each predicate is
 * evaluated in FP32 first and decided there only where T32 is farther from its threshold than the proven
 * FP32 error bound (2e-4, 4e-4 at shifted coordinates); every other cell is evaluated in FP64. Results are
 * identical in both modes. */
int dsx_set_precision(dsx_ctx *ctx, int mode);
/* Sampler::climate (standalone.hpp:224-238) at quarts (xs[k], 64, zs[k]): q[6k..6k+5] = 10000 x {T,H,C,E,D,W}. */
int dsx_climate(dsx_ctx *ctx, uint64_t seed, const int32_t *xs, const int32_t *zs, uint64_t n, int64_t *q);
/* Flood predicate `kind` (DSX_PRED_*) at quarts (xs[k], zs[k]); cut is the proxy cut (Params::pcut). */
int dsx_predicates(dsx_ctx *ctx, uint64_t seed, int kind, float cut, const int32_t *xs, const int32_t *zs, uint64_t n,
                   int8_t *out);

/* Many seeds at once: point k belongs to seeds[sidx[k]] and sits at quart (xs[k], zs[k]).
 * what = DSX_PRED_* gives the predicate in out8[k]; what = DSX_POINT_T gives T(x, 0, z) in outd[k]
 * (the hot-zone search of dsearch_desert.cpp:328-332). */
enum { DSX_POINT_T = 16, DSX_POINT_PROXYPROB = 17 };   /* 17: DesertProxy::prob (desert_proxy.hpp:25-30) in outd */
/* what = DSX_POINT_HOTTEST: (xs[k], zs[k]) is a site; the hottest of the 7 x 7 points site + (dx, dz) * 256
 * quarts, dz then dx from -3 to 3, by T(x, 0, z) with the first maximum kept (the hot-zone search of
 * dsearch_desert.cpp:328-332): outd[k] = its T, out8[k] = its index (dz + 3) * 7 + dx + 3. */
enum { DSX_POINT_HOTTEST = 18 };
/* what = DSX_POINT_BIOME: SurfaceSampler::biome (surface.hpp:79-83), the surface biome id at quart (x, 64, z)
 * from the exact climate (grid lookup, else the nearest-box search), in out8[k] read as uint8_t. */
enum { DSX_POINT_BIOME = 19 };
/* what = DSX_POINT_PROXY_C / _E: DesertProxy::octaves (desert_proxy.hpp:14-24) of continentalness over
 * octaves 0-2 / erosion over octaves 0-1 at quart (x, z), in outd (the zone features of study=1,
 * dsearch_desert.cpp:402-412). */
enum { DSX_POINT_PROXY_C = 20, DSX_POINT_PROXY_E = 21 };
int dsx_points(dsx_ctx *ctx, const uint64_t *seeds, uint64_t nseeds, const uint32_t *sidx, const int32_t *xs,
               const int32_t *zs, uint64_t n, int what, float cut, int8_t *out8, double *outd);

/* ---------------------------------------------------------------- bounded floods (S3-S4, area queries) */
/* One flood: window of (2*half+1)^2 lattice points of spacing `step` quarts centred on (qx, qz). */
typedef struct {
    uint64_t seed;
    int32_t kind;     /* DSX_PRED_* */
    float cut;        /* proxy cut (kind DSX_PRED_PROXY) */
    int32_t qx, qz;   /* centre, quarts */
    int32_t step, half;
    int32_t conn8;    /* 1: also step to diagonal neighbours */
    int32_t stop_cells; /* > 0: the flood may stop once it has counted this many cells; its cells are then
                         * >= stop_cells and its other outputs and masks unspecified (a threshold test
                         * such as the hot-zone pre-flood's, dsearch_desert.cpp:337-339, needs no more).
                         * 0: complete flood. At most 2^23. */
} dsx_flood_job;

/* standalone::DenseArea (dense_flood.hpp:9-13). */
typedef struct {
    uint64_t cells;   /* cells with predicate 2 reached from the centre */
    uint64_t sampled; /* cells whose predicate was evaluated */
    int32_t minx, maxx, minz, maxz; /* quart bounds of the counted cells, starting at the centre */
    int32_t edge;     /* the flood reached the window edge */
    int32_t border;   /* the flood touched the world border (|x| or |z| > 7,499,996) */
} dsx_area;

/* A batch of floods, evaluated together on the device. Replaces dense_flood_bridged
 * (dense_flood.hpp:52-86) at every call site: hot zones (dsearch_desert.cpp:337, 352), proxy floods
 * (:387), exact floods (:512, :521) and tools/dflood.cpp:20. Results are identical to the reference's
 * depth-first flood (they do not depend on visiting order). Windows: 2*half+1 <= DSX_FLOOD_MAX_WIDTH
 * lattice points per side (so one flood's window holds at most the engine's 2^21 tiles of 32 x 32) and
 * half*step <= 10^6 quarts; otherwise DSX_E_ARG. DSX_E_CAPACITY if the batch needs more tiles or queue
 * slots than the engine holds (split the batch). A batch whose tiled engine reaches its time limit (a queue
 * item not taken for 5 s) runs once more on a reset engine; DSX_E_HIP if it does again. Both events are
 * counted in dsx_flood_stats (time_retries, capacity_resets). */
enum { DSX_FLOOD_MAX_WIDTH = 46335 };
int dsx_flood(dsx_ctx *ctx, const dsx_flood_job *jobs, uint64_t n, dsx_area *out);

/* An evaluation hint for flood `job` of a dsx_flood_hinted batch: the 8x8 sub-blocks `blocks` (bit bz*4 + bx,
 * 16 bits) of its 32x32 window tile (tx, tz) are evaluated up front, in parallel with every other hinted
 * tile, instead of when the flood first touches them. Hints change only which cells are evaluated and when:
 * the results are those of dsx_flood (a flood's cells, sampled cells, bounds and flags follow the cells it
 * touches, never the evaluated ones). Hints for floods the small-window engine runs are ignored. */
typedef struct {
    uint32_t job;
    int32_t tx, tz;
    uint32_t blocks;
} dsx_flood_hint;
/* dsx_flood with evaluation hints (no counterpart in the reference): the pipeline hints each step-8 exact
 * flood with the proxy component that selected it (dsearch_desert.cpp:498-512) and each step-4 flood with
 * the region of the step-8 flood it refines (:520-523), so the exact floods' predicate is evaluated in one
 * parallel pass rather than along their frontiers. Tile coordinates are relative to the job's window
 * corner; hints outside the window are ignored; a hint's job must be < n (else DSX_E_ARG). At most 2^24
 * hints. */
int dsx_flood_hinted(dsx_ctx *ctx, const dsx_flood_job *jobs, uint64_t n, const dsx_flood_hint *hints, uint64_t nh,
                     dsx_area *out);

/* The counted cells (state 3 in dense_flood.hpp's window) of the floods of the last dsx_flood call, as
 * 32x32 tile masks: window cell (i, j) = (tx*32 + c, tz*32 + r) is counted iff bit c of rows[r].
 * Only tiles with at least one counted cell; any order. The hot-zone and proxy stages need them for
 * their covered sets (dsearch_desert.cpp:342-346, 358-363, 388-395). */
typedef struct {
    uint32_t job;
    int32_t tx, tz;
    uint32_t rows[32];
} dsx_tile_mask;
int dsx_flood_masks(dsx_ctx *ctx, dsx_tile_mask *out, uint64_t cap, uint64_t *n_out);

/* Flood-engine statistics, process-wide over every context (no counterpart in the reference, which
 * times its stages with wall clocks, dsearch_desert.cpp:590-593). Counted per dsx_flood call that
 * succeeds. "Sampled" cells are those the reference's flood evaluates (DenseArea::sampled), split by
 * cost class: hot = the predicate's temperature test passed, so its longer branch ran. */
typedef struct {
    uint64_t calls, rounds;        /* dsx_flood calls; tile work items the tiled engine processed */
    uint64_t floods[4];            /* floods run by the tiled engine, by kind (DSX_PRED_*) */
    uint64_t tiles[4];             /* 32x32 tiles it allocated, by kind */
    uint64_t evaluated[4];         /* cells whose predicate it evaluated, by kind */
    uint64_t sampled_cold[4];      /* sampled cells of those floods, by kind and cost class */
    uint64_t sampled_hot[4];
    uint64_t small_floods, small_sampled; /* floods of the small-window engine and their sampled cells */
    double eval_ms;                /* device time of the tiled engine's kernel (HIP events; profiling on) */
    uint64_t eval_launches;        /* its launches timed */
    /* tiled-engine floods by kind and size class (tiles per flood <= 1, 4, 16, 64, 256, more):
     * floods, tiles and sampled cells */
    uint64_t size_floods[4][6], size_tiles[4][6], size_sampled[4][6];
    uint64_t eval_rounds;          /* tile evaluation rounds of the tiled engine */
    uint64_t eval_needed;          /* 8x8 sub-blocks those rounds needed (speculative ones not counted) */
    /* with profiling on (dsx_profile): workgroup time of the tiled engine in 100 MHz ticks, summed over
     * workgroups: waiting for a queue item, evaluating sub-blocks (octave loads included), loading octaves
     * into LDS, and from start to exit */
    uint64_t wg_wait, wg_eval, wg_load, wg_life;
    /* batches that failed and were cleaned up, counted whether or not the call then succeeded:
     * time_retries = batches whose tiled engine reached its time limit (a queue item not taken for 5 s) and
     * were run again on a reset engine; capacity_resets = batches that exhausted the tile pool or the work
     * queue (the call returns DSX_E_CAPACITY). Both are 0 in a healthy run. */
    uint64_t time_retries, capacity_resets;
} dsx_flood_stats;
/* 1: bracket every launch of the tiled engine's kernel with HIP events on its stream (eval_ms). */
int dsx_profile(int on);
int dsx_flood_stats_get(dsx_flood_stats *out, int reset);

/* ---------------------------------------------------------------- the whole per-seed search */
/* struct Params of the reference (dsearch_desert.cpp:50-78), same names and defaults. */
typedef struct {
    double max_phase, amp, amp_margin, sigma, sigmac, wc, we, t, ta;
    int32_t max_sites;
    double pcut, hot_min;
    int32_t multi, exact_n;
    double exact_min, hot_pre, csv_min;
    int32_t nms;
    double refine_mb, verify_mb;
    int32_t study, gpu, gpucheck, gscan, gscancheck, gate;
    double gate_phase, gate_amp, gate_pthr;
    int32_t from_stdin;
    int32_t workers; /* pipeline workers, each with its own stream and device buffers (0: default 16) */
    int32_t prepass; /* 1: FP32 pre-pass + FP64 confirm in the floods and predicates (dsx_set_precision) */
    int32_t mcverify; /* 1: verify every CANDIDATE in-process with the exact step-1 area query (half 20000,
                       * tools/dflood.cpp) and append "<UTC time> <run> <CANDIDATE line> => 1:4 Size: N |
                       * Blocks^2 Size: M" to PREFIX.verified: the line deploy/gpipe.sh:7-11 writes with
                       * MC-Measure and scripts/log_finds.py:13 parses */
} dsx_search_params;
void dsx_search_defaults(dsx_search_params *p);

enum {
    DSX_SEARCH_ENUM = 0, /* seeds splitmix64(start+k), k < count, CPU-default gate (phase, amplitude_sub, maxta+maxtb) */
    DSX_SEARCH_GATE = 1, /* GPU gate over [start, start+count) (gate=1: gate_phase, gate_pthr, gate_amp) */
    DSX_SEARCH_LIST = 2  /* explicit gated seeds (stdin=1) */
};
typedef struct {
    uint64_t done, passed, scanned, sites;
    uint64_t hot_zones, big_zones; /* hot-zone floods run and those reaching hot_min (dsearch_desert.cpp:348, 369) */
    /* worker-seconds of the pipeline workers during the call, per stage (the reference's per-stage core-seconds,
     * dsearch_desert.cpp:552-554, 590-593, for this pipeline's stages; process-wide if searches overlap):
     * tables = tables, blur and scan on the device; scan = the host's bucket reduction, sort and NMS; sites =
     * hot zones, proxy and exact floods per seed; of these, proxy = proxy-flood rounds, hot = hot-zone starts,
     * gpu = waiting for device batches (points and floods) */
    double sec_tables, sec_scan, sec_sites, sec_proxy, sec_hot, sec_gpu;
} dsx_search_counts;

/* The per-seed pipeline of dsearch_desert.cpp:173-557 (gate -> tables/scan -> sites -> hot zones ->
 * proxy floods -> exact floods) with every numeric stage on the GPU. Appends rows to PREFIX.csv and
 * PREFIX.seeds.csv in the reference's formats (opened in append mode; the caller writes the headers),
 * with mcverify=1 the verified CANDIDATE lines to PREFIX.verified, and with study=1 (no exact floods)
 * the per-zone feature rows of dsearch_desert.cpp:401-420 to PREFIX.zones.csv (append mode, header by
 * the caller),
 * prints CANDIDATE / best lines to stderr as the reference does, and accumulates counts.
 * Rows of one seed are contiguous; seeds appear in batch order. */
int dsx_search(dsx_ctx *ctx, const dsx_search_params *p, int mode, uint64_t start, uint64_t count, const uint64_t *seeds,
               uint64_t nseeds, const char *prefix, dsx_search_counts *counts);

/* Stop (on = 1) or allow (on = 0) the searches of this process, as the reference's SIGINT / SIGTERM handler
 * does (dsearch_desert.cpp:40-41, 151, 586-598): running dsx_search calls start no new seeds, finish and write
 * the seeds already in flight, and return DSX_E_INTERRUPTED with counts of what they wrote; later calls return
 * it at once until dsx_interrupt(0). Async-signal-safe (one atomic store): call it from a signal handler. In
 * the CPU-gate mode (DSX_SEARCH_ENUM) the seed-index chunk in flight (2^22 indices) is finished. */
int dsx_interrupt(int on);

/* ---------------------------------------------------------------- area queries (tools/dflood.cpp) */
typedef struct {
    uint64_t seed;
    int32_t x, z;   /* block coordinates; the query quart is (x >> 2, z >> 2) (dflood.cpp:14) */
    int32_t half;   /* window half-size in quarts: (2*half+1)^2 cells of step 1 (dflood.cpp: 20000) */
    int32_t pad_;
} dsx_area_query;

/* Exact area queries, batched: for each query, desert_state at the query quart (dflood.cpp:17); a
 * start that is not desert (state != 2) gives cells = 0, sampled = 1 and no flood; otherwise the
 * step-1, 4-connected, river-bridged flood dense_flood_bridged(desert_state, qx, qz, 1, half)
 * (dflood.cpp:20). blocks^2 = 16 * cells (MC-Measure's "Blocks^2 Size"). */
int dsx_area_queries(dsx_ctx *ctx, const dsx_area_query *q, uint64_t n, dsx_area *out);

/* Synthetic area-query workload (no counterpart in the reference): query j has seed splitmix64(j);
 * then, from one xoroshiro128++ seeded like standalone::Random(gen_seed): X = bounded(60000001) - 3e7,
 * Z = bounded(60000001) - 3e7, u = uniform(), cap = floor(10^(3 + 4u)) cells, half =
 * floor((sqrt(cap) - 1) / 2). caps may be null. */
void dsx_area_workload(uint64_t gen_seed, uint64_t n, dsx_area_query *out, uint64_t *caps);

/* ---------------------------------------------------------------- proxy-table Monte Carlo (tools/dfunc.cpp) */
/* P(desert-passable | low-octave C, E) histograms at hot points, the input of the proxy table
 * (tools/dfunc.cpp:26-75, Makefile:58-59 `make dtable`). Seeds mix(7k + 99991), k < nseeds, that pass the
 * temperature phase filter (sum |fy - 1/2| <= 0.36); per seed, tries (x, z) = (bounded(15e6) - 7.5e6, ...)
 * from Random(seed ^ 0x1234) until pts hot points (T at the shifted position >= 0.55) or 200 pts tries.
 * hist[((set * 40 + ci) * 40 + ei) * 3 + {0, 1, 2}] = {hot points, desert, desert or bridgeable river},
 * set 0: octave 0 of C and E, set 1: C octaves 0-2, E octaves 0-1; bins of 0.05 from -0.8 (C), -1 (E). */
enum { DSX_DFUNC_CELLS = 2 * 40 * 40 * 3 };
int dsx_dfunc(dsx_ctx *ctx, uint32_t nseeds, uint32_t pts, uint64_t *hist);

/* ---------------------------------------------------------------- request server
 * A request loop on device-resident state (no counterpart in the reference, which runs one search per
 * process): seed-range scans (the dsx_search pipeline with the GPU gate) and area queries are
 * submitted at any time from any thread; worker threads, each owning a context (stream, tables,
 * flood tiles), pick the oldest pending work: a group of gated seeds of a scan, or every pending area
 * query that fits the worker's flood tiles (continuous batching), so requests that arrive while others
 * run are admitted as workers and memory free up and overlap on the device. Results do not depend on
 * the mix: a request's output equals that of running it alone. */
typedef struct dsx_server dsx_server;
typedef struct {
    int32_t workers;          /* worker contexts; 0: 16 */
    int32_t area_batch;       /* most area queries per flood batch; 0: 64 (a burst is shared by all workers) */
    uint64_t hbm_budget;      /* device bytes the worker contexts may hold (fewer workers if needed); 0: no limit */
    dsx_search_params params; /* pipeline parameters of scan requests */
} dsx_server_config;
typedef struct {
    uint64_t id;
    int32_t kind;                     /* 0 scan, 1 area query */
    int32_t status;                   /* DSX_OK or the error code of the request */
    double t_submit, t_start, t_done; /* seconds on the server's clock */
    dsx_area area;                    /* area query result */
    dsx_search_counts counts;         /* scan counts (done, passed, scanned, sites) */
    uint64_t csv_bytes, seeds_bytes;  /* scan rows: PREFIX.csv and PREFIX.seeds.csv text sizes */
} dsx_result;
int dsx_server_create(int device, const dsx_server_config *cfg, dsx_server **out);
/* Finishes the requests in flight, then stops the workers; their contexts (and their device memory) are
 * kept for the next server on the device (DSX_SERVER_POOL=0: freed here). */
int dsx_server_destroy(dsx_server *s);
/* No counterpart in the reference (serving only): frees the contexts that destroyed servers left for reuse
 * on the device. */
int dsx_server_pool_clear(int device);
int dsx_server_workers(const dsx_server *s);
/* GPU-gate scan of seeds splitmix64(k), k in [start, start + count). */
int dsx_submit_scan(dsx_server *s, uint64_t start, uint64_t count, uint64_t *id);
int dsx_submit_area(dsx_server *s, const dsx_area_query *q, uint64_t *id);
int dsx_wait(dsx_server *s, uint64_t id, dsx_result *r);   /* blocks until the request is done */
int dsx_poll(dsx_server *s, uint64_t id, dsx_result *r);   /* DSX_E_BUSY while it runs */
/* Rows of a finished scan (CSV text without headers); either buffer may be null. */
int dsx_result_text(dsx_server *s, uint64_t id, char *csv, uint64_t csv_cap, char *seeds, uint64_t seeds_cap);
/* Forget a finished request. */
int dsx_release(dsx_server *s, uint64_t id);

#ifdef __cplusplus
}
#endif
#endif /* DSX_H */
