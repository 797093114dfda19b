"""FP64 climate sampler and flood engine on the GPU against the CPU oracle (bitwise / exact), and the
exact step-1 area at verified finds against MC-Measure's cell counts."""
import concurrent.futures as cf
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, FloodOut, load_package

pytestmark = pytest.mark.gpu

FINDS = [json.loads(l) for l in open(os.path.join(GOLDEN, "desert_finds.jsonl"))]
FINDS = [r for r in FINDS if "x" in r and "cells" in r]
SEEDS = [int(r["seed"]) % (1 << 64) for r in FINDS[:3]] + [12345]


def u64(s):
    return int(s) % (1 << 64)


@pytest.mark.parametrize("ch", range(6))
def test_noise_bitexact(engine, oracle, ch):
    rng = np.random.default_rng(ch)
    for seed in SEEDS[:2]:
        xyz = np.column_stack([rng.uniform(-8e6, 8e6, 300), np.zeros(300), rng.uniform(-8e6, 8e6, 300)])
        xyz[:100, 1] = rng.uniform(-8e6, 8e6, 100)          # y != 0 branch (Shift's second sample)
        xyz[100:150] = np.round(xyz[100:150])                # integer quarts
        got = engine.noise(seed, ch, xyz)
        want = np.array([oracle.or_noise(seed, ch, *p) for p in xyz])
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_climate_vector_bitexact(engine, oracle):
    r = FINDS[0]
    seed = u64(r["seed"])
    rng = np.random.default_rng(1)
    xs = (r["x"] >> 2) + rng.integers(-3000, 3000, 400)
    zs = (r["z"] >> 2) + rng.integers(-3000, 3000, 400)
    got = engine.climate(seed, xs, zs)
    for k in range(len(xs)):
        q = np.zeros(6, np.int64)
        oracle.or_climate(seed, int(xs[k]), int(zs[k]), q.ctypes.data)
        assert np.array_equal(got[k], q), k


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_predicates_exact(engine, oracle, kind):
    """Predicates around verified deserts (many desert / river / hot cells) and at random places."""
    rng = np.random.default_rng(kind)
    for r in FINDS[:3]:
        seed = u64(r["seed"])
        xs = np.concatenate([(r["x"] >> 2) + rng.integers(-2500, 2500, 6000), rng.integers(-7_500_000, 7_500_000, 1000)])
        zs = np.concatenate([(r["z"] >> 2) + rng.integers(-2500, 2500, 6000), rng.integers(-7_500_000, 7_500_000, 1000)])
        got = engine.predicates(seed, kind, xs, zs)
        want = oracle.pred_batch(seed, kind, xs, zs)
        assert np.array_equal(got, want)
        if kind == 0:
            assert (got == 2).sum() > 500 and (got == 1).sum() > 0


def _area(a):
    return (int(a["cells"]), int(a["sampled"]), int(a["minx"]), int(a["maxx"]), int(a["minz"]), int(a["maxz"]),
            bool(a["edge"]), bool(a["border"]))


def _compare(engine, oracle, rows):
    dsx = load_package().dsx
    got = engine.flood(dsx.flood_jobs(rows))
    with cf.ThreadPoolExecutor(8) as ex:
        want = list(ex.map(lambda r: oracle.flood(*r), rows))
    for r, g, w in zip(rows, got, want):
        assert _area(g) == w, (r, _area(g), w)


def test_flood_desert_state_steps(engine, oracle):
    """Bridged desert_state floods at the verified finds, steps 8 (4- and 8-connected) and 4."""
    rows = []
    for r in FINDS[:6]:
        seed, qx, qz = u64(r["seed"]), r["x"] >> 2, r["z"] >> 2
        rows += [(seed, 0, 0.35, qx, qz, 8, 2500, 0), (seed, 0, 0.35, qx, qz, 8, 2000, 1), (seed, 0, 0.35, qx, qz, 4, 4000, 0)]
    _compare(engine, oracle, rows)


def test_flood_hot_and_proxy(engine, oracle):
    """Hot-zone floods (step 256 and 128, 4-connected) and proxy floods (step 32, 8-connected)."""
    rows = []
    for r in FINDS[:8]:
        seed, qx, qz = u64(r["seed"]), r["x"] >> 2, r["z"] >> 2
        rows += [(seed, 1, 0.35, qx, qz, 256, 125, 0), (seed, 1, 0.35, qx, qz, 128, 250, 0),
                 (seed, 2, 0.35, qx, qz, 32, 500, 1), (seed, 2, 0.5, qx + 640, qz - 320, 32, 500, 1)]
    _compare(engine, oracle, rows)


def test_flood_edges_and_borders(engine, oracle):
    """Window edge hits, world border, a start that is not passable, a start outside the world."""
    r = FINDS[0]
    seed, qx, qz = u64(r["seed"]), r["x"] >> 2, r["z"] >> 2
    rows = [(seed, 0, 0.35, qx, qz, 8, 60, 0), (seed, 0, 0.35, qx, qz, 4, 33, 1), (seed, 0, 0.35, qx, qz, 1, 0, 0),
            (seed, 1, 0.35, 7_499_000, 100, 256, 125, 0), (seed, 1, 0.35, -7_499_990, 7_499_990, 64, 200, 1),
            (seed, 1, 0.35, 7_499_998, 0, 64, 40, 0), (seed, 0, 0.35, qx + 20000, qz + 20000, 4, 500, 0),
            (seed, 2, 0.35, 0, 0, 32, 31, 1), (seed, 1, 0.35, 5, 5, 16, 16, 0)]
    _compare(engine, oracle, rows)


def test_flood_small_windows_mixed_batch(engine, oracle):
    """Windows up to 255 x 255 (single-wave floods) of every kind, mixed with larger ones in one batch."""
    rows = []
    for r in FINDS[6:14]:
        seed, qx, qz = u64(r["seed"]), r["x"] >> 2, r["z"] >> 2
        rows += [(seed, 0, 0.35, qx, qz, 8, 127, 1), (seed, 0, 0.35, qx, qz, 16, 127, 0),
                 (seed, 3, 0.35, qx, qz, 8, 100, 0), (seed, 1, 0.35, qx, qz, 256, 127, 0),
                 (seed, 1, 0.35, qx, qz, 64, 128, 0), (seed, 2, 0.35, qx, qz, 32, 127, 1),
                 (seed, 2, 0.6, qx - 96, qz + 64, 16, 90, 1), (seed, 0, 0.35, qx, qz, 2, 1, 0)]
    _compare(engine, oracle, rows)


def test_flood_stop_cells(engine, oracle):
    """stop_cells (the hot-zone pre-flood's threshold test): a flood that counts fewer cells than the stop
    is complete and equals the oracle; one that reaches it reports at least that many cells, and so does
    the oracle's complete flood. 138 is the pipeline's stop (0.8 x 180 M at step 256)."""
    dsx = load_package().dsx
    rows = []
    for r in FINDS[:10]:
        seed, qx, qz = u64(r["seed"]), r["x"] >> 2, r["z"] >> 2
        rows += [(seed, 1, 0.35, qx, qz, 256, 125, 0), (seed, 1, 0.35, qx + 3000, qz - 7000, 256, 125, 0),
                 (seed, 1, 0.35, qx - 20000, qz + 9000, 256, 125, 0), (seed, 1, 0.35, qx, qz, 64, 100, 0)]
    with cf.ThreadPoolExecutor(8) as ex:
        want = list(ex.map(lambda r: oracle.flood(*r), rows))
    # a stop never reached: every flood completes; the large hot zones leave the 64 x 64 sub-window of
    # the threshold-flood instance and are finished by the tiled engine, exactly
    jobs = dsx.flood_jobs(rows)
    jobs["stop_cells"] = 1 << 23
    got = engine.flood(jobs)
    for k, (row, w) in enumerate(zip(rows, want)):
        assert _area(got[k]) == w, ("no stop", k, row)
    assert max(w[0] for w in want) > 64 * 64   # some of them cannot fit the sub-window
    for stop in (138, 1, 40):
        jobs = dsx.flood_jobs(rows)
        jobs["stop_cells"] = stop
        got = engine.flood(jobs)
        n_capped = 0
        for k, (row, w) in enumerate(zip(rows, want)):
            if w[0] < stop:
                assert _area(got[k]) == w, (stop, k, row)
            else:
                assert int(got[k]["cells"]) >= stop, (stop, k, row)
                n_capped += int(got[k]["cells"]) < w[0]
        assert n_capped > 0, stop   # some floods really stopped early
    bad = dsx.flood_jobs(rows[:1])
    bad["stop_cells"] = -1
    with pytest.raises(dsx.DsxError):
        engine.flood(bad)


def test_hottest_point_search(engine, oracle):
    """The device search for the hottest of 7 x 7 points around each site (dsearch_desert.cpp:328-332:
    dz then dx from -3 to 3, step 256, strict greater, so the first maximum wins) equals that scan over
    T at the 49 points, and T there equals the oracle's NormalNoise."""
    dsx = load_package().dsx
    rng = np.random.default_rng(7)
    seeds = np.array(SEEDS, np.uint64)
    n = 600
    sidx = rng.integers(0, len(seeds), n).astype(np.uint32)
    xs = rng.integers(-7_400_000, 7_400_000, n).astype(np.int32)
    zs = rng.integers(-7_400_000, 7_400_000, n).astype(np.int32)
    xs[:8] = 7_499_000   # sites near the world border
    o8, od = engine.points(seeds, sidx, xs, zs, dsx.POINT_HOTTEST)
    off = np.arange(49)
    px = (xs[:, None] + (off % 7 - 3)[None, :] * 256).ravel()
    pz = (zs[:, None] + (off // 7 - 3)[None, :] * 256).ravel()
    _, t = engine.points(seeds, np.repeat(sidx, 49), px, pz, dsx.POINT_T)
    t = t.reshape(n, 49)
    for k in range(n):
        best, bi = -9.0, 0
        for q in range(49):
            if t[k, q] > best:
                best, bi = t[k, q], q
        assert (int(o8[k]), od[k]) == (bi, best), k
    for k in range(0, n, 97):   # T itself against the oracle (bitwise)
        for q in (0, 24, 48):
            w = oracle.or_noise(int(seeds[sidx[k]]), 0, float(px[k * 49 + q]), 0.0, float(pz[k * 49 + q]))
            assert t[k, q] == w


def test_flood_counted_masks(engine, oracle):
    """The counted cells each flood exports (state 3 of dense_flood.hpp's window) equal the oracle's
    window state, for every engine: small windows, proxy and hot-zone floods that run on a local
    sub-window of the one-wave engine, those that outgrow it (tiled engine), and exact floods."""
    dsx = load_package().dsx
    rows = []
    for r in FINDS[14:20]:
        seed, qx, qz = u64(r["seed"]), r["x"] >> 2, r["z"] >> 2
        rows += [(seed, 2, 0.35, qx + 2048, qz - 1024, 32, 500, 1), (seed, 2, 0.35, qx, qz, 32, 500, 1),
                 (seed, 1, 0.35, qx - 4096, qz, 128, 250, 0), (seed, 1, 0.35, qx, qz, 256, 125, 0),
                 (seed, 0, 0.35, qx, qz, 8, 300, 1), (seed, 2, 0.5, qx + 512, qz + 512, 32, 100, 1)]
    got = engine.flood(dsx.flood_jobs(rows))
    masks = engine.flood_masks()
    for k, row in enumerate(rows):
        seed, kind, cut, qx, qz, step, half, c8 = row
        n = 2 * half + 1
        st = np.zeros(n * n, np.uint8)
        out = FloodOut()
        oracle.L.or_flood(seed, kind, cut, qx, qz, step, half, c8, ctypes.byref(out), st.ctypes.data)
        want = st.reshape(n, n) == 3
        have = np.zeros((((n + 31) // 32) * 32,) * 2, bool)
        for m in masks[masks["job"] == k]:
            bits = (m["rows"][:, None] >> np.arange(32, dtype=np.uint32)[None, :]) & 1
            have[m["tz"] * 32:m["tz"] * 32 + 32, m["tx"] * 32:m["tx"] * 32 + 32] |= bits.astype(bool)
        assert np.array_equal(have[:n, :n], want), row
        assert not have[n:, :].any() and not have[:, n:].any()
        assert int(got[k]["cells"]) == int(want.sum()) == out.cells


def test_area_queries_step1_match_mcmeasure(engine):
    """The exact step-1 flood (tools/dflood.cpp: half 20000) at verified finds = MC-Measure's cells."""
    dsx = load_package().dsx
    rows = [(u64(r["seed"]), 0, 0.35, r["x"] >> 2, r["z"] >> 2, 1, 20000, 0) for r in FINDS[:4]]
    got = engine.flood(dsx.flood_jobs(rows))
    for r, g in zip(FINDS[:4], got):
        assert int(g["cells"]) == r["cells"] and not g["edge"] and not g["border"]


def test_flood_golden_step4_all_finds(engine):
    """Step-4 floods at all 220 finds in one batch against the reference dflood readings."""
    dsx = load_package().dsx
    gold = [json.loads(l) for l in open(os.path.join(GOLDEN, "dflood.jsonl"))]
    gold = [g for g in gold if g["step"] == 4]
    rows = [(u64(g["seed"]), 0, 0.35, g["x"] >> 2, g["z"] >> 2, 4, 5000, 0) for g in gold]
    got = engine.flood(dsx.flood_jobs(rows))
    assert [int(a["cells"]) for a in got] == [g["cells"] for g in gold]


def test_flood_hints_change_nothing(engine):
    """dsx_flood_hinted: evaluation hints only change which cells are evaluated and when. Step-4 floods hinted
    with their step-8 flood's region (as the pipeline hints them), plus hints that help nothing (every
    sub-block of random tiles, tiles outside the window, duplicates, hints of small-window floods) give
    results and counted-cell masks identical to dsx_flood."""
    dsx = load_package().dsx
    rng = np.random.default_rng(11)
    pts = [(u64(r["seed"]), r["x"] >> 2, r["z"] >> 2) for r in FINDS[20:26]]
    f8 = engine.flood(dsx.flood_jobs([(s, 0, 0.35, qx, qz, 8, 2000, 1) for s, qx, qz in pts]))
    m8 = engine.flood_masks()
    rows = [(s, 0, 0.35, qx, qz, 4, 4000, 0) for s, qx, qz in pts]
    rows += [(pts[0][0], 2, 0.35, pts[0][1], pts[0][2], 32, 500, 1), (pts[1][0], 1, 0.35, pts[1][1], pts[1][2], 128, 250, 0),
             (pts[2][0], 0, 0.35, pts[2][1], pts[2][2], 4, 60, 0)]   # the last fits the small-window engine
    jobs = dsx.flood_jobs(rows)
    want = engine.flood(jobs)
    wmask = engine.flood_masks()
    hints = []
    for k in range(len(pts)):   # step-4 sub-block (b, c) covers step-8 cells [4b, 4b+4) x [4c, 4c+4)
        marked = set()
        for m in m8[m8["job"] == k]:
            for r in range(32):
                bits = int(m["rows"][r])
                for g in range(8):
                    if (bits >> (4 * g)) & 0xF:
                        b, c = int(m["tx"]) * 8 + g, int(m["tz"]) * 8 + r // 4
                        marked |= {(b + db, c + dc) for db in (-1, 0, 1) for dc in (-1, 0, 1)}
        tiles = {}
        for b, c in marked:
            if 0 <= b < 1001 and 0 <= c < 1001:
                tiles[(c // 4, b // 4)] = tiles.get((c // 4, b // 4), 0) | (1 << ((c % 4) * 4 + b % 4))
        hints += [(k, tx, tz, bits) for (tz, tx), bits in tiles.items()]
    assert len(hints) > 100
    nta = [((2 * h + 1) + 31) // 32 for h in (4000,) * len(pts) + (500, 250, 60)]
    for _ in range(300):
        k = int(rng.integers(len(rows)))
        hints.append((k, int(rng.integers(-2, nta[k] + 2)), int(rng.integers(-2, nta[k] + 2)), 0xFFFF))
    hints += hints[:50]
    h = np.array(hints, dtype=dsx.FLOOD_HINT)
    got = engine.flood_hinted(jobs, h)
    gmask = engine.flood_masks()
    assert [tuple(a) for a in got] == [tuple(a) for a in want]
    key = lambda m: (int(m["job"]), int(m["tx"]), int(m["tz"]), tuple(int(x) for x in m["rows"]))
    assert sorted(map(key, gmask)) == sorted(map(key, wmask))
    assert [int(a["cells"]) for a in f8] and all(int(a["cells"]) > 0 for a in got[:len(pts)])


def test_flood_hints_self_staged_batch(engine):
    """A batch without small-window floods is staged by the tiled engine's own first workgroups, and a hinted
    tile can be queued by another workgroup than the one that stores its job: every job is stored before any
    tile is queued. 440 step-4 exact floods (301^2 windows, tiled) with hints in reverse job order (so the
    hints of the last jobs are staged by the first staging workgroup, their jobs by the second), run three
    times with the hints in a new order each time: results and counted-cell masks equal to dsx_flood's."""
    dsx = load_package().dsx
    rng = np.random.default_rng(5)
    rows = [(u64(r["seed"]), 0, 0.35, (r["x"] >> 2) + dx, (r["z"] >> 2) + dz, 4, 150, 0)
            for r in FINDS for dx, dz in ((0, 0), (37, -23))]
    assert len(rows) >= 400
    jobs = dsx.flood_jobs(rows)
    want = engine.flood(jobs)
    key = lambda m: (int(m["job"]), int(m["tx"]), int(m["tz"]), tuple(int(x) for x in m["rows"]))
    wmask = sorted(map(key, engine.flood_masks()))
    t = 150 // 32   # the start tile (150, 150) and its neighbours in the 10 x 10 tile window
    hints = [(k, t + dx, t + dz, 0xFFFF) for k in reversed(range(len(rows))) for dx in (-1, 0, 1) for dz in (-1, 0, 1)]
    for rep in range(3):
        h = np.array(hints if rep == 0 else [hints[i] for i in rng.permutation(len(hints))], dtype=dsx.FLOOD_HINT)
        got = engine.flood_hinted(jobs, h)
        assert [tuple(a) for a in got] == [tuple(a) for a in want], rep
        assert sorted(map(key, engine.flood_masks())) == wmask, rep
    assert sum(int(a["cells"]) > 0 for a in want) > 200
