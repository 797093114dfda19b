"""The whole per-seed pipeline on the GPU (dsx_search, and the dsearch_desert command line) against
the reference's own outputs: sorted CSV rows, seeds CSV and JSON counters, byte for byte."""
import hashlib
import json
import os
import subprocess
import tempfile

import pytest

from conftest import GOLDEN, REPO, load_package

pytestmark = pytest.mark.gpu


def _check(tag, prefix, counts=None, stderr=None):
    """stderr: the text the run printed; its CANDIDATE lines, sorted, must equal the reference's."""
    gold = json.load(open(os.path.join(GOLDEN, tag + ".json")))
    rows = sorted(open(prefix + ".csv").read().splitlines()[1:])
    srows = sorted(open(prefix + ".seeds.csv").read().splitlines()[1:])
    want = open(os.path.join(GOLDEN, tag + ".csv")).read().splitlines()
    bad = [(a, b) for a, b in zip(rows, want) if a != b]
    assert len(rows) == len(want) and not bad, f"{len(rows)} vs {len(want)} rows; first difference {bad[:1]}"
    assert len(srows) == gold["seeds_rows"]
    assert hashlib.sha256("\n".join(srows).encode()).hexdigest() == gold["seeds_sha256"]
    if counts is not None:
        assert list(counts) == [gold["done"], gold["passed"], gold["scanned"], gold["sites"]]
    if stderr is not None:
        cand = sorted(l for l in stderr.splitlines() if l.startswith("CANDIDATE "))
        path = os.path.join(GOLDEN, tag + ".cand")
        gcand = open(path).read().splitlines() if os.path.exists(path) else []
        assert len(gcand) == gold.get("candidates", 0)
        assert cand == gcand, f"{len(cand)} vs {len(gcand)} CANDIDATE lines"


def _seeds(name):
    return [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, name))]


def test_config1_pipeline(engine):
    """The 200,000-index workload: seeds mix(k), k < 200000, CPU-default gate semantics, production Params."""
    dsx = load_package().dsx
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "c1")
        counts = engine.search(pre, dsx.SEARCH_ENUM, 0, 200000)
        _check("c1", pre, counts)


def test_gated_list_pipeline(engine):
    """stdin=1 semantics on 92 GPU-gated seeds (exact step-8 floods included)."""
    dsx = load_package().dsx
    seeds = [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, "gated_1e12_1m.txt"))]
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "rg1m")
        counts = engine.search(pre, dsx.SEARCH_LIST, seeds=seeds)
        _check("rg1m", pre, counts)


def _check_zones(tag, prefix):
    """PREFIX.zones.csv (study=1) against the reference's sorted zone rows."""
    import gzip
    rows = sorted(open(prefix + ".zones.csv").read().splitlines()[1:])
    gold = gzip.open(os.path.join(GOLDEN, tag + ".zones.csv.gz"), "rt").read().splitlines()
    assert len(rows) == len(gold)
    bad = [(a, b) for a, b in zip(rows, gold) if a != b]
    assert not bad, f"{len(bad)} zone rows differ, first: {bad[0]}"


def test_study_zones(engine):
    """study=1 on the 92 gated seeds: no exact floods (s8 = -1 in every row), and one PREFIX.zones.csv
    row per proxy component with its step-64 hot zone and low-octave C/E census."""
    dsx = load_package().dsx
    seeds = [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, "gated_1e12_1m.txt"))]
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "st1m")
        counts = engine.search(pre, dsx.SEARCH_LIST, seeds=seeds, params=engine.search_params(study=1))
        _check("st1m", pre, counts)
        _check_zones("st1m", pre)


def test_cli_study(engine):
    """dsearch_desert stdin=1 study=1: the zones file gets the reference's header and rows; an existing
    zones file is refused like the other outputs."""
    exe = os.path.join(REPO, "desert-search_amd", "dsearch_desert")
    lines = open(os.path.join(GOLDEN, "gated_1e12_1m.txt")).read()
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "st")
        r = subprocess.run([exe, "0", "0", pre, "8", "stdin=1", "study=1"], input=lines, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        assert open(pre + ".zones.csv").readline().startswith("seed,phT,amp,phE0,phE1,phC0,phC1,rank,x,z,")
        _check_zones("st1m", pre)
        os.remove(pre + ".csv")
        os.remove(pre + ".seeds.csv")
        r2 = subprocess.run([exe, "0", "0", pre, "8", "stdin=1", "study=1"], input="", capture_output=True, text=True,
                            timeout=60)
        assert r2.returncode == 2 and "exist?" in r2.stderr


@pytest.mark.parametrize("prepass", [0, 1])
def test_exact_branch_pipeline(engine, capfd, prepass):
    """The exact-flood driver (dsearch_desert.cpp:498-531) with exact_min=0 refine=0 verify=0 on the 92
    gated seeds: every step-8 result is refined by the 4-connected step-4 flood (s4 and the ORed edge flag
    in the rows) and printed as a CANDIDATE; rows and CANDIDATE lines equal the reference's."""
    dsx = load_package().dsx
    p = engine.search_params(exact_min=0, refine_mb=0, verify_mb=0, prepass=prepass)
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "rgx")
        capfd.readouterr()
        counts = engine.search(pre, dsx.SEARCH_LIST, seeds=_seeds("gated_1e12_1m.txt"), params=p)
        _check("rgx", pre, counts, stderr=capfd.readouterr().err)


def test_production_candidates_pipeline(engine, capfd):
    """Production Params (stdin=1) on the seeds of the verified finds and two seeds whose CANDIDATE lines a
    GPU run printed, with the FP32 pre-pass: step-4 floods at the production thresholds and the
    CANDIDATE list equal the reference's."""
    dsx = load_package().dsx
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "fx")
        capfd.readouterr()
        counts = engine.search(pre, dsx.SEARCH_LIST, seeds=_seeds("fx_seeds.txt"),
                               params=engine.search_params(prepass=1))
        _check("fx", pre, counts, stderr=capfd.readouterr().err)


def test_cli_production_candidates(engine):
    """The same through the dsearch_desert command line (stdin=1), plain FP64: rows, seeds CSV, JSON
    counters and the CANDIDATE lines on stderr."""
    exe = os.path.join(REPO, "desert-search_amd", "dsearch_desert")
    lines = open(os.path.join(GOLDEN, "fx_seeds.txt")).read()
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "fx")
        r = subprocess.run([exe, "0", "0", pre, "8", "stdin=1"], input=lines, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        js = json.load(open(pre + ".json"))
        _check("fx", pre, (js["done"], js["passed"], js["scanned"], js["sites"]), stderr=r.stderr)


def test_cli_alternative_paths(engine):
    """The paths the defaults do not take, through the command line: the gate launched stage by stage
    (DSX_GATE_GRAPH=0, no fast pre-gate decisions), accumulate and publish as two kernels
    (DSX_FLOOD_ACC_PUBLISH=0), exact floods without evaluation hints (DSX_HINT_EXACT=0) and the library's
    default launch settings (DSX_TUNED=0). gate=1 over [10^12, 10^12 + 2^20) gives the reference's rows
    and counters (rg1m), and stdin=1 on the production-Params seeds its rows and CANDIDATE lines (fx)."""
    exe = os.path.join(REPO, "desert-search_amd", "dsearch_desert")
    env = dict(os.environ, DSX_GATE_GRAPH="0", DSX_FLOOD_ACC_PUBLISH="0", DSX_HINT_EXACT="0", DSX_TUNED="0")
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "g")
        r = subprocess.run([exe, str(10 ** 12), str(1 << 20), pre, "8", "gpu=1", "gate=1"], capture_output=True,
                           text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stderr
        rows = sorted(open(pre + ".csv").read().splitlines()[1:])
        assert rows == open(os.path.join(GOLDEN, "rg1m.csv")).read().splitlines()
        js, gold = json.load(open(pre + ".json")), json.load(open(os.path.join(GOLDEN, "rg1m.json")))
        assert [js[k] for k in ("done", "passed", "scanned", "sites")] == \
            [gold[k] for k in ("done", "passed", "scanned", "sites")]
        pre = os.path.join(td, "fx")
        r = subprocess.run([exe, "0", "0", pre, "8", "stdin=1"], input=open(os.path.join(GOLDEN, "fx_seeds.txt")).read(),
                           capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stderr
        js = json.load(open(pre + ".json"))
        _check("fx", pre, (js["done"], js["passed"], js["scanned"], js["sites"]), stderr=r.stderr)


def test_cli_validation_switches(engine):
    """gscancheck=N and gpucheck=N (dsearch_desert.cpp:228-242, 289-296, 445-450): the first seeds' tables and
    scan recomputed on the CPU agree with the GPU (maxima, every bucket's best score), and the CPU
    DesertProxy agrees with the GPU proxy at every 7th cell of the first big hot zones' grids; the outputs
    are unchanged."""
    import re
    exe = os.path.join(REPO, "desert-search_amd", "dsearch_desert")
    lines = open(os.path.join(GOLDEN, "gated_1e12_1m.txt")).read()
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "chk")
        r = subprocess.run([exe, "0", "0", pre, "8", "stdin=1", "gpu=1", "gscan=1", "gscancheck=3", "gpucheck=2"],
                           input=lines, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        js = json.load(open(pre + ".json"))
        _check("rg1m", pre, (js["done"], js["passed"], js["scanned"], js["sites"]))
    gs = re.findall(r"gscancheck seed (-?\d+): maxta ([\d.-]+)/([\d.-]+) maxtb ([\d.-]+)/([\d.-]+), buckets cpu (\d+) "
                    r"gpu (\d+), mismatched (\d+), raw (\d+)", r.stderr)
    assert len(gs) == 3, r.stderr[-2000:]
    for seed, a1, a2, b1, b2, nc, ng, miss, raw in gs:
        assert a1 == a2 and b1 == b2 and nc == ng and miss == "0" and int(nc) > 0, (seed, a1, a2, b1, b2, nc, ng, miss)
    gp = re.findall(r"gpucheck zone (\d+)x(\d+): (\d+)/(\d+) cells differ", r.stderr)
    assert len(gp) == 2
    for nx, nz, diff, tot in gp:
        assert diff == "0" and int(tot) > 100
    bad = subprocess.run([exe, "0", "10", os.path.join(td, "x"), "1", "gscancheck=1"], capture_output=True, text=True)
    assert bad.returncode == 2 and "gscan=1 needs gpu=1" in bad.stderr


def test_cli_interrupt(engine):
    """SIGINT mid-run (dsearch_desert.cpp:40-41, 586-598): the seeds in flight are finished and written, the JSON
    is written and the exit status is 0; every row is whole and belongs to a seed of the seeds CSV, and the
    counters equal the rows written (gate=1: done = passed = scanned = gated seeds processed)."""
    import signal
    import time
    exe = os.path.join(REPO, "desert-search_amd", "dsearch_desert")
    total = 1 << 34
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "int")
        p = subprocess.Popen([exe, str(10 ** 12), str(total), pre, "8", "gpu=1", "gate=1"], stderr=subprocess.PIPE,
                             stdout=subprocess.DEVNULL, text=True)
        t0 = time.time()
        while time.time() - t0 < 120 and p.poll() is None:
            if os.path.exists(pre + ".seeds.csv") and os.path.getsize(pre + ".seeds.csv") > 4096:
                break
            time.sleep(0.2)
        p.send_signal(signal.SIGINT)
        err = p.communicate(timeout=120)[1]
        assert p.returncode == 0, err[-2000:]
        js = json.load(open(pre + ".json"))
        srows = open(pre + ".seeds.csv").read().splitlines()[1:]
        rows = open(pre + ".csv").read().splitlines()[1:]
    assert 0 < js["done"] == js["passed"] == js["scanned"] == len(srows) < total // 20000
    assert all(len(r.split(",")) == 8 for r in srows)
    seeds = {r.split(",")[0] for r in srows}
    assert all(len(r.split(",")) == 17 and r.split(",")[0] in seeds for r in rows)
    assert ("done: %d seeds (%d passed)" % (js["done"], js["passed"])) in err
    assert js["cpu_tables"] > 0 and js["cpu_sites"] > 0


def test_3m_pipeline(engine):
    dsx = load_package().dsx
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "r3m")
        counts = engine.search(pre, dsx.SEARCH_ENUM, 0, 3000000)
        _check("r3m", pre, counts)


def test_gate_mode_equals_gated_list(engine):
    """gate=1 over [1e12, 1e12 + 2^20) gives the same rows as stdin=1 on the gate's survivors."""
    dsx = load_package().dsx
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "g")
        counts = engine.search(pre, dsx.SEARCH_GATE, 10 ** 12, 1 << 20)
        _check("rg1m", pre, counts)


def test_cli_config1(engine):
    """The dsearch_desert command line: same outputs and JSON counters as the reference binary."""
    exe = os.path.join(REPO, "desert-search_amd", "dsearch_desert")
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "cli")
        r = subprocess.run([exe, "0", "200000", pre, "8"], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert "done: 200000 seeds (3891 passed)" in r.stderr
        js = json.load(open(pre + ".json"))
        _check("c1", pre, (js["done"], js["passed"], js["scanned"], js["sites"]))
        r2 = subprocess.run([exe, "0", "10", pre, "8"], capture_output=True, text=True, timeout=60)
        assert r2.returncode == 2 and "exist?" in r2.stderr
        r3 = subprocess.run([exe, "0", "10", pre + "x", "8", "bogus=1"], capture_output=True, text=True, timeout=60)
        assert r3.returncode == 2 and "unknown key bogus" in r3.stderr


def test_mcverify_candidates(engine):
    """mcverify=1: every CANDIDATE gets the exact step-1 area (half 20000) in PREFIX.verified, in the line
    format scripts/log_finds.py:13 parses. Thresholds are lowered so that the exact floods run and every
    step-4 result becomes a CANDIDATE; the areas equal dsx_area_queries at the same points."""
    import re
    import numpy as np
    dsx = load_package().dsx
    seeds = [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, "gated_1e12_1m.txt"))][:24]
    p = engine.search_params(exact_min=0, refine_mb=0, verify_mb=0, mcverify=1)
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "v")
        engine.search(pre, dsx.SEARCH_LIST, seeds=seeds, params=p)
        lines = open(pre + ".verified").read().splitlines()
    pat = re.compile(r'^(\S+) (\S+) CANDIDATE step4=([\d.]+)M seed=(-?\d+) x=(-?\d+) z=(-?\d+).*1:4 Size: (\d+) \| '
                     r'Blocks\^2 Size: (\d+)')
    assert len(lines) >= 10
    q = np.zeros(len(lines), dsx.AREA_QUERY)
    want = []
    for k, l in enumerate(lines):
        m = pat.search(l)
        assert m, l
        assert m.group(2) == "v"
        cells, b2 = int(m.group(7)), int(m.group(8))
        assert b2 == 16 * cells
        q[k] = (int(m.group(4)) % (1 << 64), int(m.group(5)), int(m.group(6)), 20000, 0)
        want.append(cells)
    got = engine.area_queries(q)
    assert [int(c) for c in got["cells"]] == want
