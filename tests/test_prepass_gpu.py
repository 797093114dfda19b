"""FP32 pre-pass + FP64 confirm (dsx_set_precision 1) against plain FP64: the FP32 noise stays within the
proven error bound, and floods, predicates and the whole pipeline give identical results."""
import json
import os
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, load_package

pytestmark = pytest.mark.gpu

FINDS = [json.loads(l) for l in open(os.path.join(GOLDEN, "desert_finds.jsonl"))]
FINDS = [r for r in FINDS if "x" in r and "cells" in r]


def u64(s):
    return int(s) % (1 << 64)


def test_fp32_noise_within_bound(engine):
    """|T32 - T64| <= 2e-4 (the bound the pre-pass decides with) at random world points; the other
    channels' FP32 errors are reported alongside (the pre-pass decides only on T)."""
    rng = np.random.default_rng(11)
    worst = {}
    for seed in [u64(r["seed"]) for r in FINDS[:4]]:
        pts = np.zeros((20000, 3))
        pts[:, 0] = rng.integers(-7_500_000, 7_500_000, len(pts))
        pts[:, 2] = rng.integers(-7_500_000, 7_500_000, len(pts))
        for ch in range(6):
            d = np.abs(engine.noise_f32(seed, ch, pts) - engine.noise(seed, ch, pts)).max()
            worst[ch] = max(worst.get(ch, 0.0), float(d))
    print("FP32 max abs error by channel (T, H, C, E, shift, W):", worst)
    assert worst[0] < 2e-4
    assert all(v < 1e-3 for v in worst.values())


def _floods(engine, rows):
    dsx = load_package().dsx
    out = engine.flood(dsx.flood_jobs(rows))
    masks = engine.flood_masks()
    order = np.lexsort((masks["tz"], masks["tx"], masks["job"]))
    return out, masks[order]


def test_floods_identical_in_both_modes(engine):
    rows = []
    for r in FINDS[20:28]:
        seed, qx, qz = u64(r["seed"]), r["x"] >> 2, r["z"] >> 2
        rows += [(seed, 0, 0.35, qx, qz, 8, 2000, 1), (seed, 0, 0.35, qx, qz, 4, 600, 0), (seed, 3, 0.35, qx, qz, 8, 300, 0),
                 (seed, 1, 0.35, qx, qz, 256, 125, 0), (seed, 1, 0.35, qx, qz, 128, 250, 0),
                 (seed, 2, 0.35, qx, qz, 32, 500, 1), (seed, 2, 0.5, qx + 640, qz, 32, 500, 1),
                 (seed, 0, 0.35, qx, qz, 1, 100, 0)]
    try:
        engine.set_precision(0)
        a64, m64 = _floods(engine, rows)
        engine.set_precision(1)
        a32, m32 = _floods(engine, rows)
    finally:
        engine.set_precision(0)
    assert np.array_equal(a64, a32)
    assert np.array_equal(m64, m32)


def test_predicates_identical_in_both_modes(engine):
    rng = np.random.default_rng(5)
    r = FINDS[30]
    seed = u64(r["seed"])
    xs = (r["x"] >> 2) + rng.integers(-3000, 3000, 50000)
    zs = (r["z"] >> 2) + rng.integers(-3000, 3000, 50000)
    try:
        for kind in range(4):
            engine.set_precision(0)
            p64 = engine.predicates(seed, kind, xs, zs)
            engine.set_precision(1)
            p32 = engine.predicates(seed, kind, xs, zs)
            assert np.array_equal(p64, p32), kind
    finally:
        engine.set_precision(0)


def test_pipeline_with_prepass_equals_reference(engine):
    """stdin=1 semantics on the 92 GPU-gated seeds with prepass=1: the reference's rows byte for byte."""
    import hashlib
    dsx = load_package().dsx
    seeds = [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, "gated_1e12_1m.txt"))]
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "p")
        counts = engine.search(pre, dsx.SEARCH_LIST, seeds=seeds, params=engine.search_params(prepass=1))
        gold = json.load(open(os.path.join(GOLDEN, "rg1m.json")))
        rows = sorted(open(pre + ".csv").read().splitlines()[1:])
        srows = sorted(open(pre + ".seeds.csv").read().splitlines()[1:])
        assert rows == open(os.path.join(GOLDEN, "rg1m.csv")).read().splitlines()
        assert hashlib.sha256("\n".join(srows).encode()).hexdigest() == gold["seeds_sha256"]
        assert list(counts) == [gold["done"], gold["passed"], gold["scanned"], gold["sites"]]
