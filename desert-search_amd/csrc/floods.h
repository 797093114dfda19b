// Tiled frontier flood engine: shared declarations.
#pragma once
#include "climate.h"

namespace dsx {

constexpr int FT = 32;   // tile edge in lattice cells; one 32-bit mask per tile row

// One bounded flood (dense_flood.hpp:52-86): window of (2*half+1)^2 lattice points, spacing `step`
// quarts, centred on (qx, qz); pred kind (PredKind), proxy cut, 4- or 8-connectivity.
struct FloodJob {
    int32_t seed;   // index into the batch's ClimSeed array
    int32_t kind;
    float cut;
    int32_t qx, qz, step, half;
    int32_t conn8;  // bit 0: 8-connectivity; bits 8-31: stop after this many counted cells (0: never)
};
struct FloodRes {
    unsigned long long cells, sampled;
    int minx, maxx, minz, maxz;
    int edge, border;
    unsigned long long hot;     // sampled cells that passed the predicate's temperature test
    unsigned long long tiles;   // tiles the engine allocated for this flood (0: small-window engine)
    unsigned long long evald;   // cells it evaluated (8x8 sub-blocks x 64)
};
// A 32x32 block of one flood's window: window cell (i, j) = (tx*32 + c, tz*32 + r) <-> bit c of row r.
struct Tile {
    uint32_t job;
    int32_t tx, tz;
    uint32_t qstate;        // 0 idle, 1 queued or being processed, 2 being processed and touched again
    uint32_t slot1;         // hash slot + 1 of this tile's entry (0: not in the table)
    uint32_t evald;         // 8x8 sub-blocks evaluated so far (bit bz*4 + bx)
    uint32_t hint;          // sub-blocks to evaluate at the tile's first visit, touched or not (a flood hint)
    uint32_t pad_;
    uint32_t pass[FT];      // predicate != 0 (inside window and world)
    uint32_t cnt[FT];       // predicate == 2
    uint32_t reach[FT];     // cells the flood has entered (pushed)
    uint32_t in[FT];        // cells touched from a neighbouring tile (and the start cell)
    uint32_t hot[FT];       // the predicate's temperature test passed (its costlier branch ran)
};

struct TileMask {
    uint32_t job;
    int32_t tx, tz;
    uint32_t rows[FT];
};

// Device state of the tiled engine: the tile pool, the (job, tx, tz) -> tile hash, and the work
// queue of tiles to process (indices; Q_EMPTY where unused, which every call leaves behind).
struct FloodBufs {
    Tile *tiles;
    uint32_t *ntiles;             // allocated tiles
    uint32_t tile_cap;
    unsigned long long *hkeys;    // hash: (job, tx, tz) -> tile index, packed with the key
    uint32_t hmask;               // table size - 1
    uint32_t *wq;                 // work queue slots
    uint32_t wq_cap;
    uint32_t *ctl;                // this batch's control words (zero when it starts)
    uint32_t *ctl_next;           // the control words of the batch two after this one (zeroed by this one)
    uint32_t seq;                 // batch sequence number (for the time-limit diagnostics)
    struct TileMask *h_masks;     // when set: exported masks are also written here as they are made (pinned,
    uint32_t h_mcap;              // at most h_mcap), for batches whose publishing copies no masks
};

// ctl words: queue head/tail, pending items, overflow flags, exported masks, items processed, floods
// handed over by the small-window engine, allocated tiles (B.ntiles points here); then what the first
// workgroup that reached the time limit saw (for the error message): a flag, its slot, the queue head,
// tail and pending count, and the slot's content; evaluation rounds of the tiled engine and the
// sub-blocks they needed (before speculative ones are added); with profiling on, the workgroups' time
// (100 MHz ticks) waiting for queue items, evaluating sub-blocks, loading octaves into LDS, and alive;
// then the batch number (plus one) the stage kernel wrote, the one the tiled engine's first workgroup
// was given, the queue tail and the stage's batch number that workgroup read when it started; the
// accumulate kernel's workgroups that have finished; when the tiled engine stages the batch itself, the
// tickets its workgroups drew on arrival, the staging workgroups that have finished, and those that have
// stored their jobs
constexpr int Q_HEAD = 0, Q_TAIL = 1, Q_PENDING = 2, FLOOD_OVF = 3, FLOOD_MASKS = 4, Q_ITEMS = 5, HANDED = 6,
              CTL_NTILES = 7, TIME_SEEN = 8, TIME_SLOT = 9, TIME_HEAD = 10, TIME_TAIL = 11, TIME_PENDING = 12,
              TIME_ITEM = 13, EVAL_ROUNDS = 14, EVAL_NEED = 15, PROF_WAIT = 16, PROF_EVAL = 17, PROF_LOAD = 18,
              PROF_LIFE = 19, STAGE_SEQ = 20, RUN_SEQ = 21, RUN_TAIL0 = 22, RUN_STAGE_SEQ = 23, ACC_DONE = 24,
              STAGE_TICKET = 25, STAGE_DONE = 26, STAGE_JOBS = 27, CTL_WORDS = 28;
constexpr uint32_t Q_EMPTY = 0xffffffffu;
// FLOOD_OVF bits: 1 work queue full, 2 tile pool or hash table full, 4 time limit reached
constexpr uint32_t OVF_QUEUE = 1, OVF_TILES = 2, OVF_TIME = 4;

// Evaluation hint of a tiled flood: the 8x8 sub-blocks `blocks` (bit bz*4 + bx) of the tile (tx, tz) of
// job `job` are evaluated when the tile is first visited, whether or not the flood has touched them yet.
struct FloodHint {
    uint32_t job;
    int32_t tx, tz;
    uint32_t blocks;
};
// First kernel of a batch of n floods, reading the host's pinned staging directly: jobs and job lists
// (layout: n jobs, n small-window indices, n tiled-engine indices, then nh hints) to the device, every
// result initialised (an empty area at the centre), the start tile of each of the n_big tiled floods
// touched and queued, and every hinted tile allocated, given its hint and queued.
void launch_flood_stage(hipStream_t st, const uint8_t *h_up, uint32_t n, uint32_t n_small, uint32_t n_big, uint32_t nh,
                        uint8_t *d_up, FloodRes *res, FloodBufs B);
// A batch staged by the tiled engine's own first workgroups instead of k_flood_stage (batches without
// small-window floods): the stage kernel's arguments; nst = the number of staging workgroups (set by
// launch_flood_run; 0 = the batch is already staged).
struct StageArgs {
    const uint8_t *h_up;
    uint8_t *d_up;
    FloodRes *res;
    uint32_t n, n_big, nh, nst;
};
// The persistent flood kernel: `blocks` workgroups take queued tiles until no work is pending. A tile
// is evaluated lazily, by 8x8 sub-blocks the flood touches; every processing grows its reached set to
// a fixpoint and queues the neighbours it touched. ev0/ev1 (optional): HIP events around it.
// pre: 1 = FP32 pre-pass on the predicates' temperature test (identical results); light: 1 when every
// tiled job is a hot-zone or proxy flood (selects the instance without the biome predicates);
// stage (optional): stage the batch in this kernel (no k_flood_stage launch before it).
void launch_flood_run(hipStream_t st, const FloodJob *jobs, const ClimSeed *cs, const SurfGrid &g, FloodBufs B,
                      uint32_t blocks, int pre, int light, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                      int tprof = 0, const StageArgs *stage = nullptr);
// Reduce every allocated tile (count read on the device) into res, export counted masks, and return
// the tiles and hash slots to the clean state.
void launch_flood_acc(hipStream_t st, const FloodJob *jobs, FloodBufs B, FloodRes *res, struct TileMask *masks,
                      uint32_t mcap, uint32_t blocks);
// Copy the results, the control words and up to hmcap exported masks into host memory, so that one
// stream synchronisation ends the batch, and zero the next batch's control words.
void launch_flood_publish(hipStream_t st, const FloodRes *res, uint32_t n, FloodBufs B, const struct TileMask *masks,
                          FloodRes *h_res, uint32_t *h_ctl, struct TileMask *h_masks, uint32_t hmcap);
// Both in one kernel: the accumulation's last workgroup to finish publishes.
void launch_flood_acc_publish(hipStream_t st, const FloodJob *jobs, FloodBufs B, FloodRes *res, struct TileMask *masks,
                              uint32_t mcap, uint32_t blocks, uint32_t n, FloodRes *h_res, uint32_t *h_ctl,
                              struct TileMask *h_masks, uint32_t hmcap);
// One wave per flood with its state in LDS over a local sub-window of at most LDS_MAX_N^2 points
// around the start; a flood that outgrows it is handed to the tiled engine (its result reset and its
// start tile queued). cls: LDS_HOT (the T test, any window), LDS_PROXY (proxy predicate, any window),
// LDS_FULL (any kind; the host sends only windows that fit whole).
constexpr int LDS_MAX_N = 255;
// LDS_HOT_SMALL / LDS_PROXY_SMALL: 64 x 64 sub-window (floods that leave it are handed to the tiled engine)
enum { LDS_HOT = 0, LDS_PROXY = 1, LDS_FULL = 2, LDS_HOT_SMALL = 3, LDS_PROXY_SMALL = 4, LDS_NCLS = 5 };
void launch_flood_lds(hipStream_t st, int cls, const FloodJob *jobs, const uint32_t *list, uint32_t n, const ClimSeed *cs,
                      const SurfGrid &g, FloodRes *res, TileMask *masks, uint32_t mcap, FloodBufs B, int pre);


}  // namespace dsx
