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
    int32_t qx, qz, step, half, conn8;
};
struct FloodRes {
    unsigned long long cells, sampled;
    int minx, maxx, minz, maxz;
    int edge, border;
    unsigned long long hot;     // sampled cells that passed the predicate's temperature test
    unsigned long long tiles;   // tiles the engine allocated for this flood (0: small-window engine)
};
// A 32x32 block of one flood's window: window cell (i, j) = (tx*32 + c, tz*32 + r) <-> bit c of row r.
struct Tile {
    uint32_t job;
    int32_t tx, tz;
    uint32_t queued;        // 1 while listed for the next propagation round
    uint32_t slot1;         // hash slot + 1 of this tile's entry (0: not in the table)
    uint32_t pad_[3];
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

struct FloodBufs {
    Tile *tiles;
    uint32_t *ntiles;             // allocated tiles
    uint32_t tile_cap;
    unsigned long long *hkeys;    // hash: (job, tx, tz) -> tile index, packed with the key
    uint32_t hmask;               // table size - 1
    uint32_t *lists[6];           // eval / prop lists of rounds r with r % 3 = 0, 1, 2
    uint32_t *lcount;             // [0..5] their lengths; [FLOOD_OVF] overflow flags; [FLOOD_MASKS] exported masks
    uint32_t list_cap;
};

constexpr int FLOOD_OVF = 6, FLOOD_MASKS = 7;

// jl[0..n): indices of the jobs to start (their hash keys and result slots use the original index);
// the start tiles go to the lists of round 0
void launch_flood_start(hipStream_t st, const FloodJob *jobs, const uint32_t *jl, uint32_t n, FloodBufs B, uint32_t *eval,
                        uint32_t *prop);
// Round `round`: evaluate the new tiles of its eval list, then propagate the tiles of its prop list
// into the next round's lists. `width`: the tiles the host expects per list (a grid hint: any grid
// covers every listed tile).
constexpr uint32_t EVAL_GRID = 512, PROP_GRID = 2048;
// ev0/ev1 (optional): HIP events recorded when the evaluation kernel starts and ends.
void launch_flood_round(hipStream_t st, const FloodJob *jobs, const ClimSeed *cs, const SurfGrid &g, FloodBufs B, int round,
                        uint32_t width, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// Reduce, export the counted masks, and return every touched tile and hash slot to the zero / empty
// state the next call expects (no pool-wide memsets).
void launch_flood_acc(hipStream_t st, const FloodJob *jobs, FloodBufs B, uint32_t ntiles, FloodRes *res,
                      struct TileMask *masks, uint32_t *mcount, uint32_t mcap);
// res[k] = empty area at the centre, for k in jl[0..n) (jl null: k in [0, n))
void launch_flood_resinit(hipStream_t st, const FloodJob *jobs, const uint32_t *jl, uint32_t n, FloodRes *res);
// One wave per small-window flood (2*half+1 <= 255), state in LDS; hot: predicate is the T test.
constexpr int LDS_MAX_N = 255;
void launch_flood_lds(hipStream_t st, bool hot, const FloodJob *jobs, const uint32_t *list, uint32_t n, const ClimSeed *cs,
                      const SurfGrid &g, FloodRes *res, TileMask *masks, uint32_t *mcount, uint32_t mcap);

}  // namespace dsx
