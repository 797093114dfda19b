// Proxy-table Monte Carlo on libdsx.so with the command line and output of the reference's
// tools/dfunc.cpp:1-76:
//   dfunc SEEDS POINTS THREADS > dfunc.txt        (THREADS is accepted and unused: the GPU samples)
// Rows "set ci ei total desert passable" for every histogram cell with samples, set-major then C bin then
// E bin (set 0: C and E octave 0; set 1: C octaves 0-2, E octaves 0-1). gen_dtable.py turns them into the
// 40 x 40 proxy table.
#include "../../include/dsx.h"
#include "hwq.h"
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

int main(int argc, char **argv) {
    dsx_set_hw_queues();
    if (argc < 4) {
        std::fprintf(stderr, "dfunc SEEDS POINTS THREADS\n");
        return 2;
    }
    const int seeds = std::stoi(argv[1]), pts = std::stoi(argv[2]);
    if (seeds < 0 || pts < 0) {
        std::fprintf(stderr, "dfunc: SEEDS and POINTS must be >= 0\n");
        return 2;
    }
    dsx_ctx *ctx = nullptr;
    int rc = dsx_create(0, &ctx);
    std::vector<uint64_t> H(DSX_DFUNC_CELLS);
    if (!rc) rc = dsx_dfunc(ctx, uint32_t(seeds), uint32_t(pts), H.data());
    if (rc) {
        std::fprintf(stderr, "dfunc: %s %s\n", dsx_strerror(rc), dsx_last_error());
        return 1;
    }
    dsx_destroy(ctx);
    for (int set = 0; set < 2; set++)
        for (int i = 0; i < 40; i++)
            for (int j = 0; j < 40; j++) {
                const uint64_t *cell = &H[size_t((set * 40 + i) * 40 + j) * 3];
                if (cell[0])
                    std::printf("%d %d %d %llu %llu %llu\n", set, i, j, (unsigned long long)cell[0],
                                (unsigned long long)cell[1], (unsigned long long)cell[2]);
            }
    return 0;
}
