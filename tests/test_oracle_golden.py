"""The CPU oracle against the reference's own outputs (golden fixtures written by
tests/golden/make_goldens.py from the reference binaries) and the MC-Measure-verified finds."""
import concurrent.futures as cf
import ctypes
import hashlib
import json
import os
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN


def _run(oracle, tmp, tag, start=0, count=0, seeds=None, exact=None):
    """exact = (exact_min, refine, verify) overrides the exact-flood thresholds."""
    pre = os.path.join(tmp, tag)
    cnt = np.zeros(4, np.uint64)
    sec = ctypes.c_double()
    par = np.array([0.10, 0.70, 0.01] + list(exact or (75, 70, 80)), np.float64)
    if seeds is not None:
        arr = np.ascontiguousarray(seeds, dtype=np.uint64)
        rc = oracle.or_run(0, 0, arr.ctypes.data, len(arr), 1, 8, pre.encode(), cnt.ctypes.data, ctypes.byref(sec),
                           par.ctypes.data, len(par))
    else:
        rc = oracle.or_run(start, count, None, 0, 0, 8, pre.encode(), cnt.ctypes.data, ctypes.byref(sec),
                           par.ctypes.data, len(par))
    assert rc == 0
    rows = sorted(open(pre + ".csv").read().splitlines()[1:])
    srows = sorted(open(pre + ".seeds.csv").read().splitlines()[1:])
    cand = sorted(open(pre + ".cand").read().splitlines())
    return rows, srows, [int(c) for c in cnt], cand


def _check(tag, rows, srows, cnt, cand=None):
    gold = json.load(open(os.path.join(GOLDEN, tag + ".json")))
    grows = open(os.path.join(GOLDEN, tag + ".csv")).read().splitlines()
    assert rows == grows
    assert len(srows) == gold["seeds_rows"]
    assert hashlib.sha256("\n".join(srows).encode()).hexdigest() == gold["seeds_sha256"]
    assert cnt == [gold["done"], gold["passed"], gold["scanned"], gold["sites"]]
    if cand is not None:
        path = os.path.join(GOLDEN, tag + ".cand")
        want = open(path).read().splitlines() if os.path.exists(path) else []
        assert len(want) == gold.get("candidates", 0)
        assert cand == want


def test_config1_matches_reference(oracle):
    """The 200,000-index workload (seeds mix(k), k < 200000, production Params): CSV, seeds CSV and counters."""
    with tempfile.TemporaryDirectory() as td:
        rows, srows, cnt, cand = _run(oracle, td, "c1", 0, 200000)
    _check("c1", rows, srows, cnt, cand)
    assert srows == open(os.path.join(GOLDEN, "c1.seeds.csv")).read().splitlines()


def test_gated_list_matches_reference(oracle):
    """stdin=1 mode on seeds gated with the GPU-gate rules (92 seeds, k in [1e12, 1e12 + 2^20))."""
    seeds = [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, "gated_1e12_1m.txt"))]
    with tempfile.TemporaryDirectory() as td:
        rows, srows, cnt, cand = _run(oracle, td, "rg1m", seeds=seeds)
    _check("rg1m", rows, srows, cnt, cand)


def test_3m_range_matches_reference(oracle):
    """3M seed indices: 392 scanned seeds including exact step-8 floods."""
    with tempfile.TemporaryDirectory() as td:
        rows, srows, cnt, cand = _run(oracle, td, "r3m", 0, 3000000)
    _check("r3m", rows, srows, cnt, cand)


def test_exact_branch_matches_reference(oracle):
    """The exact-flood driver (dsearch_desert.cpp:498-531) with exact_min=0 refine=0 verify=0 on the 92
    gated seeds: every step-8 result gets the step-4 flood (s4, edge flags ORed) and a CANDIDATE line."""
    seeds = [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, "gated_1e12_1m.txt"))]
    with tempfile.TemporaryDirectory() as td:
        rows, srows, cnt, cand = _run(oracle, td, "rgx", seeds=seeds, exact=(0, 0, 0))
    assert sum(1 for r in rows if float(r.split(",")[14]) > 0) > 100
    _check("rgx", rows, srows, cnt, cand)


def test_production_candidates_match_reference(oracle):
    """Production Params on the seeds of the verified finds (stdin=1): step-4 floods at the production
    thresholds and the CANDIDATE list, byte for byte."""
    seeds = [int(l) % (1 << 64) for l in open(os.path.join(GOLDEN, "fx_seeds.txt"))]
    with tempfile.TemporaryDirectory() as td:
        rows, srows, cnt, cand = _run(oracle, td, "fx", seeds=seeds)
    assert len(cand) > 150
    _check("fx", rows, srows, cnt, cand)


def test_fixture_rows_consistent():
    """The verified-finds fixture: blocks^2 = 16 * cells, CANDIDATE coordinates 4q + 2."""
    n = 0
    for line in open(os.path.join(GOLDEN, "desert_finds.jsonl")):
        r = json.loads(line)
        if "cells" not in r or "x" not in r:
            continue
        n += 1
        assert r["cells"] * 16 == r["blocks2"]
        assert r["x"] % 4 == 2 and r["z"] % 4 == 2
    assert n == 220


class FloodOut(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_uint64), ("sampled", ctypes.c_uint64), ("minx", ctypes.c_int32),
                ("maxx", ctypes.c_int32), ("minz", ctypes.c_int32), ("maxz", ctypes.c_int32),
                ("edge", ctypes.c_int32), ("border", ctypes.c_int32)]


def _flood(oracle, seed, x, z, step):
    out = FloodOut()
    oracle.or_flood(seed % (1 << 64), 0, 0.35, x >> 2, z >> 2, step, 20000 // step, 0, ctypes.byref(out), None)
    return out


@pytest.mark.parametrize("step,every", [(8, 1), (4, 6)])
def test_dflood_matches_reference(oracle, step, every):
    """dflood SEED X Z STEP (bridged desert_state flood, 4-connected, half = 20000/step) at the finds."""
    gold = [json.loads(l) for l in open(os.path.join(GOLDEN, "dflood.jsonl"))]
    gold = [g for g in gold if g["step"] == step][::every]
    with cf.ThreadPoolExecutor(8) as ex:
        res = list(ex.map(lambda g: _flood(oracle, g["seed"], g["x"], g["z"], step), gold))
    for g, r in zip(gold, res):
        assert r.cells == g["cells"], g
        assert bool(r.edge) == g["edge"]


def test_dflood_step1_matches_mcmeasure(oracle):
    """The exact step-1 flood at one verified find equals MC-Measure's cell count."""
    finds = [json.loads(l) for l in open(os.path.join(GOLDEN, "desert_finds.jsonl"))]
    r = finds[0]
    out = _flood(oracle, int(r["seed"]), r["x"], r["z"], 1)
    assert out.cells == r["cells"] and not out.edge and not out.border


def test_biomemap_matches_reference(oracle):
    """The biome-map restatement against tools/biomemap.cpp's own output (biome ids, desert component)."""
    d = np.load(os.path.join(GOLDEN, "biomemap.npz"))
    for key in [k for k in d.files if k.startswith("map_")]:
        seed, x, z, half, step = (int(a) for a in d["args_" + key[4:]])
        n, m = oracle.biomemap(seed % (1 << 64), x, z, half, step)
        assert n == d[key].shape[0]
        assert np.array_equal(m, d[key]), key
    n, _ = oracle.biomemap(1, 0, 0, 400, 16)   # the reference tool: "centre is not desert", exit 1
    assert n == -1


def test_gen_dtable_on_reference_rows():
    """gen_dtable.py on the reference's dfunc rows: P = passable / total of set 1 with 3 decimals, 0 below
    20 samples, in the layout gen_tables.py reads (its parser recovers the same 1600 values)."""
    import importlib.util
    import re
    from conftest import REPO
    spec = importlib.util.spec_from_file_location("gen_dtable", os.path.join(REPO, "desert-search_amd", "data",
                                                                             "gen_dtable.py"))
    gd = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gd)
    rows = np.loadtxt(os.path.join(GOLDEN, "dfunc_2000_2000.txt"), dtype=np.int64)
    tot = np.zeros((40, 40), np.int64)
    ok = np.zeros((40, 40), np.int64)
    r1 = rows[rows[:, 0] == 1]
    np.add.at(tot, (r1[:, 1], r1[:, 2]), r1[:, 3])
    np.add.at(ok, (r1[:, 1], r1[:, 2]), r1[:, 5])
    text = gd.header(open(os.path.join(GOLDEN, "dfunc_2000_2000.txt")).read().splitlines())
    assert "measured on %.1fM hot points" % (tot.sum() / 1e6) in text
    assert "inline constexpr float desert_pass_prob[40][40] = {" in text
    body = text[text.index("= {") + 3:]
    vals = np.array([float(v) for v in re.findall(r"(-?\d+\.\d+)f", body)]).reshape(40, 40)
    want = np.where(tot >= 20, ok / np.maximum(tot, 1), 0.0)
    assert np.all(np.abs(vals - want) <= 0.0005 + 1e-12)
    assert np.all(vals[tot < 20] == 0)
