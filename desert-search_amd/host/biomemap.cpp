// Surface biome map around a block on libdsx.so, with the command line and output of the reference's
// tools/biomemap.cpp:1-45:
//   biomemap SEED X Z [HALF_BLOCKS=16000] [STEP_BLOCKS=16] > map.bin
// Output: int32 n = 2 * HALF / STEP + 1, then n * n bytes, row-major with z outer: the surface biome id at
// quart ((X - HALF + i * STEP) >> 2, (Z - HALF + j * STEP) >> 2), or 255 on the centre's desert component
// (4-connected on the map grid, bridged rivers crossed but not marked: desert_state 2 marks, 1 crosses).
// Biome ids and desert states come from the GPU (dsx_points: DSX_POINT_BIOME, DSX_PRED_DESERT_STATE); the
// component walk over the map's bytes stays on the host.
#include "../../include/dsx.h"
#include "hwq.h"
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static int fail(const char *what, int rc) {
    std::fprintf(stderr, "%s: %s %s\n", what, dsx_strerror(rc), dsx_last_error());
    return 1;
}

int main(int argc, char **argv) {
    dsx_set_hw_queues();
    if (argc < 4) {
        std::fprintf(stderr, "biomemap SEED X Z [HALF_BLOCKS] [STEP_BLOCKS]\n");
        return 2;
    }
    const uint64_t seed = std::stoull(argv[1]);
    const int X = std::stoi(argv[2]), Z = std::stoi(argv[3]);
    const int half = argc > 4 ? std::stoi(argv[4]) : 16000, step = argc > 5 ? std::stoi(argv[5]) : 16;
    if (half < 0 || step <= 0) {
        std::fprintf(stderr, "biomemap: HALF_BLOCKS >= 0 and STEP_BLOCKS > 0\n");
        return 2;
    }
    const int n = 2 * half / step + 1;
    const size_t nn = size_t(n) * n;
    dsx_ctx *ctx = nullptr;
    int rc = dsx_create(0, &ctx);
    if (rc) return fail("dsx_create", rc);
    std::vector<uint8_t> b(nn), st(nn);
    // rows in chunks of about 2^20 points: one dsx_points call per chunk and quantity
    const int rows = std::max(1, (1 << 20) / n);
    std::vector<int32_t> xs, zs;
    std::vector<uint32_t> sidx;
    for (int j0 = 0; j0 < n; j0 += rows) {
        const int j1 = std::min(n, j0 + rows);
        const size_t m = size_t(j1 - j0) * n;
        xs.resize(m);
        zs.resize(m);
        sidx.assign(m, 0u);
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < n; i++) {
                const size_t k = size_t(j - j0) * n + i;
                xs[k] = (X - half + i * step) >> 2;
                zs[k] = (Z - half + j * step) >> 2;
            }
        int8_t *ob = reinterpret_cast<int8_t *>(b.data() + size_t(j0) * n);
        int8_t *os = reinterpret_cast<int8_t *>(st.data() + size_t(j0) * n);
        if ((rc = dsx_points(ctx, &seed, 1, sidx.data(), xs.data(), zs.data(), m, DSX_POINT_BIOME, 0.f, ob, nullptr)))
            return fail("dsx_points", rc);
        if ((rc = dsx_points(ctx, &seed, 1, sidx.data(), xs.data(), zs.data(), m, DSX_PRED_DESERT_STATE, 0.f, os,
                             nullptr)))
            return fail("dsx_points", rc);
    }
    dsx_destroy(ctx);
    const size_t c0 = size_t(n / 2) * n + n / 2;
    if (st[c0] != 2) {
        std::fprintf(stderr, "centre is not desert\n");
        return 1;
    }
    // component of the centre over cells with state 1 or 2; state-2 cells are marked 255
    std::vector<uint8_t> seen(nn);
    std::vector<size_t> stack{c0};
    seen[c0] = 1;
    while (!stack.empty()) {
        const size_t c = stack.back();
        stack.pop_back();
        if (st[c] == 2) b[c] = 255;
        const int ci = int(c % n), cj = int(c / n);
        const int nb[4][2] = {{ci + 1, cj}, {ci - 1, cj}, {ci, cj + 1}, {ci, cj - 1}};
        for (const auto &p : nb) {
            if (p[0] < 0 || p[1] < 0 || p[0] >= n || p[1] >= n) continue;
            const size_t d = size_t(p[1]) * n + p[0];
            if (!seen[d] && st[d]) {
                seen[d] = 1;
                stack.push_back(d);
            }
        }
    }
    std::fwrite(&n, 4, 1, stdout);
    std::fwrite(b.data(), 1, b.size(), stdout);
    return 0;
}
