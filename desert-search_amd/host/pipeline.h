// Host pipeline internals shared by dsx_search (search.cpp) and the request server (serve.cpp).
#pragma once
#include "../../include/dsx.h"
#include <memory>
#include <string>
#include <vector>

namespace dsxh {

struct SeedOut;

// Consecutive gated seeds processed together by one worker context: tables + scan, sites, hot zones,
// proxy and exact floods (dsearch_desert.cpp:173-557).
struct Group {
    std::vector<uint64_t> seeds;
    std::vector<std::unique_ptr<SeedOut>> outs;
    int rc = DSX_OK;
    std::string err;
    bool done = false;
    bool scanned = false;   // tables, scan and sites done (outs filled) by a scan worker
    Group();
    ~Group();
};
int run_group(dsx_ctx *w, const dsx_search_params &P, int mode, Group &g);

// The group's output in the reference's formats: PREFIX.csv rows, PREFIX.seeds.csv rows, and the
// stderr lines (CANDIDATE; "best" lines when `best` is given: it holds the running best).
struct GroupText {
    std::string csv, seeds, log, zones;
    std::vector<std::string> verified;   // verified CANDIDATE lines (mcverify=1), without time and run
    uint64_t scanned = 0, sites = 0, hot_zones = 0, big_zones = 0;
};
void format_group(const Group &g, const dsx_search_params &P, GroupText &t, double *best);

int worker_contexts(dsx_ctx *ctx, int n, std::vector<dsx_ctx *> &out);
// The last dsx_flood call's exported masks in place (dsx_flood_masks without the copy): na of them at a,
// then nb at b; valid until the context's next flood call.
int flood_masks_view(dsx_ctx *c, const dsx_tile_mask **a, uint64_t *na, const dsx_tile_mask **b, uint64_t *nb);
uint64_t env_u64(const char *name, uint64_t def);

}  // namespace dsxh
