"""Area queries (tools/dflood.cpp semantics) and the request server: exact against the CPU oracle and
the reference's own outputs, and the same result for every request alone and in a mix."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_package

pytestmark = pytest.mark.gpu

FINDS = [json.loads(l) for l in open(os.path.join(GOLDEN, "desert_finds.jsonl"))]
FINDS = [r for r in FINDS if "x" in r and "cells" in r]


def u64(s):
    return int(s) % (1 << 64)


def _queries(rows):
    dsx = load_package().dsx
    q = np.zeros(len(rows), dsx.AREA_QUERY)
    for k, r in enumerate(rows):
        q[k] = tuple(r) + (0,)
    return q


def _t(a):
    return (int(a["cells"]), int(a["sampled"]), int(a["minx"]), int(a["maxx"]), int(a["minz"]), int(a["maxz"]),
            bool(a["edge"]), bool(a["border"]))


def test_area_queries_vs_oracle(engine, oracle):
    """Step-1 windows at finds (desert starts) and next to them, against the oracle's flood; a start
    that is not desert has no flood (dflood.cpp:17)."""
    rows = []
    for r in FINDS[:6]:
        s = u64(r["seed"])
        rows += [(s, r["x"], r["z"], 150), (s, r["x"] + 4000, r["z"] - 2000, 90), (s, r["x"], r["z"], 300)]
    got = engine.area_queries(_queries(rows))
    n_flood = 0
    for r, g in zip(rows, got):
        seed, X, Z, half = r
        qx, qz = X >> 2, Z >> 2
        st = int(oracle.pred_batch(seed, 0, [qx], [qz])[0])
        if st != 2:
            assert _t(g) == (0, 1, qx, qx, qz, qz, False, False), r
        else:
            n_flood += 1
            assert _t(g) == oracle.flood(seed, 0, 0.35, qx, qz, 1, half, 0), r
    assert n_flood >= 6


def test_area_queries_match_mcmeasure(engine):
    """half = 20000 (dflood.cpp:20) at verified finds: MC-Measure's 1:4 cell counts."""
    rows = [(u64(r["seed"]), r["x"], r["z"], 20000) for r in FINDS[4:8]]
    got = engine.area_queries(_queries(rows))
    assert [int(a["cells"]) for a in got] == [r["cells"] for r in FINDS[4:8]]
    assert not any(got["edge"]) and not any(got["border"])


def test_area_workload_batch_matches_per_query(engine):
    """The synthetic workload's first 512 queries in one batch equal the same queries one by one
    (each window decided alone), and the large windows agree with small batches."""
    dsx = load_package().dsx
    q, caps = dsx.area_workload(512)
    big = engine.area_queries(q)
    assert (big["cells"] > 0).sum() >= 5
    for k in np.flatnonzero(big["cells"] > 0)[:12]:
        one = engine.area_queries(q[k:k + 1])
        assert _t(one[0]) == _t(big[k])
    for lo in range(0, 512, 128):
        part = engine.area_queries(q[lo:lo + 128])
        assert [_t(a) for a in part] == [_t(a) for a in big[lo:lo + 128]]


def test_area_workload_4096_vs_oracle(engine, oracle):
    """The whole synthetic area workload (4096 queries: seed splitmix64(j), X and Z uniform in +-3e7,
    windows from caps log-uniform over 1e3..1e7 cells) in one dsx_area_queries batch, every field against
    the oracle: the step-1, 4-connected, river-bridged flood dense_flood_bridged(desert_state, qx, qz, 1,
    half) where the start is desert (dflood.cpp:17-23, dense_flood.hpp:52-86), else no flood."""
    import concurrent.futures as cf
    import os as _os
    dsx = load_package().dsx
    q, caps = dsx.area_workload(4096)
    got = engine.area_queries(q)
    qx = (q["x"].astype(np.int64) >> 2).astype(np.int32)
    qz = (q["z"].astype(np.int64) >> 2).astype(np.int32)
    state = np.array([oracle.pred_batch(int(q[j]["seed"]), 0, [qx[j]], [qz[j]])[0] for j in range(len(q))])
    starts = np.flatnonzero(state == 2)
    for j in np.flatnonzero(state != 2):
        assert _t(got[j]) == (0, 1, qx[j], qx[j], qz[j], qz[j], False, False), j
    with cf.ThreadPoolExecutor(min(16, _os.cpu_count() or 1)) as ex:
        want = list(ex.map(lambda j: oracle.flood(int(q[j]["seed"]), 0, 0.35, int(qx[j]), int(qz[j]), 1,
                                                  int(q[j]["half"]), 0), starts))
    bad = [(int(j), _t(got[j]), w) for j, w in zip(starts, want) if _t(got[j]) != w]
    assert not bad, f"{len(bad)} of {len(starts)} floods differ, first {bad[0]}"
    assert len(starts) >= 50 and sum(w[6] for w in want) >= 20   # floods that reach their window's edge
    test_area_workload_4096_vs_oracle.result = [_t(a) for a in got]


@pytest.fixture(scope="module")
def server():
    dsx = load_package().dsx
    s = dsx.Server(0, workers=6)
    yield s
    s.close()


def test_server_scan_equals_reference(server):
    """A scan request over [1e12, 1e12 + 2^20) gives the reference's rows (gate=1 | stdin=1)."""
    rid = server.submit_scan(10 ** 12, 1 << 20)
    r = server.wait(rid)
    assert r.status == 0
    csv, seeds = server.text(rid)
    assert sorted(csv.splitlines()) == open(os.path.join(GOLDEN, "rg1m.csv")).read().splitlines()
    gold = json.load(open(os.path.join(GOLDEN, "rg1m.json")))
    assert [r.counts.done, r.counts.passed, r.counts.scanned, r.counts.sites] == \
        [gold["done"], gold["passed"], gold["scanned"], gold["sites"]]
    server.release(rid)


def test_server_area_workload_4096(server, engine):
    """The 4096 workload queries submitted to the server at once (continuous batching across its workers)
    equal the single-batch results, which the oracle pins (test above)."""
    dsx = load_package().dsx
    q, caps = dsx.area_workload(4096)
    want = getattr(test_area_workload_4096_vs_oracle, "result", None) or [_t(a) for a in engine.area_queries(q)]
    ids = [server.submit_area(int(r["seed"]), int(r["x"]), int(r["z"]), int(r["half"])) for r in q]
    for k, rid in enumerate(ids):
        r = server.wait(rid)
        assert r.status == 0
        a = r.area
        assert (a.cells, a.sampled, a.minx, a.maxx, a.minz, a.maxz, bool(a.edge), bool(a.border)) == want[k], k
        server.release(rid)


def _free_device_bytes():
    import ctypes
    L = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert L.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_server_hbm_budget():
    """dsx_server_config.hbm_budget admits as many workers as the budget holds of the first worker's measured
    device footprint (at least one, at most the workers asked for): a 1-byte budget gives one worker, no
    budget all of them, and a budget of 2.5 one-worker servers (the gate context plus one worker, so at least
    one worker's footprint each) two or three. A budget-limited server still gives the reference's rows."""
    dsx = load_package().dsx
    dsx.load().dsx_server_pool_clear(0)   # the contexts of servers closed before: s1 starts from fresh ones
    free0 = _free_device_bytes()
    s1 = dsx.Server(0, workers=6, hbm_budget=1)
    try:
        assert s1.workers == 1
        one = free0 - _free_device_bytes()
        assert one > 1 << 28
        rid = s1.submit_scan(10 ** 12, 1 << 20)
        r = s1.wait(rid)
        assert r.status == 0
        csv, _ = s1.text(rid)
        assert sorted(csv.splitlines()) == open(os.path.join(GOLDEN, "rg1m.csv")).read().splitlines()
        s1.release(rid)
    finally:
        s1.close()
    s2 = dsx.Server(0, workers=6, hbm_budget=int(2.5 * one))
    try:
        print(f"one-worker server {one / 2 ** 30:.2f} GiB; budget {2.5 * one / 2 ** 30:.2f} GiB -> {s2.workers} workers",
              flush=True)
        assert 2 <= s2.workers <= 3
    finally:
        s2.close()
    s3 = dsx.Server(0, workers=3, hbm_budget=0)
    try:
        assert s3.workers == 3
    finally:
        s3.close()


def test_server_pool_reuse():
    """A destroyed server's contexts serve the next server on the device (no new device memory), with the same
    results; dsx_server_pool_clear returns their memory."""
    dsx = load_package().dsx
    lib = dsx.load()
    q, _ = dsx.area_workload(256)
    def run():
        ids = [s.submit_area(int(r["seed"]), int(r["x"]), int(r["z"]), int(r["half"])) for r in q]
        out = []
        for rid in ids:
            r = s.wait(rid)
            assert r.status == 0
            a = r.area
            out.append((a.cells, a.sampled, a.minx, a.maxx, a.minz, a.maxz, a.edge, a.border))
            s.release(rid)
        return out
    lib.dsx_server_pool_clear(0)
    s = dsx.Server(0, workers=3)
    try:
        first = run()
    finally:
        s.close()
    free_pooled = _free_device_bytes()
    s = dsx.Server(0, workers=3)
    try:
        assert _free_device_bytes() >= free_pooled - (64 << 20)   # reused: no new contexts
        assert run() == first
    finally:
        s.close()
    assert lib.dsx_server_pool_clear(0) == 0
    freed = _free_device_bytes() - free_pooled
    print(f"pool of 4 contexts held {freed / 2 ** 30:.2f} GiB", flush=True)
    assert freed > 1 << 28


def test_server_alone_equals_mix(server):
    """Every request gives the same result alone and when submitted in a mix with others in flight."""
    dsx = load_package().dsx
    q, caps = dsx.area_workload(96)
    scans = [(10 ** 12 + (1 << 20), 1 << 19), (5 * 10 ** 11, 3 << 18), (12345, 1 << 18)]
    alone_scan = []
    for st, n in scans:
        rid = server.submit_scan(st, n)
        r = server.wait(rid)
        assert r.status == 0
        alone_scan.append((server.text(rid), (r.counts.done, r.counts.scanned, r.counts.sites)))
        server.release(rid)
    alone_area = []
    for k in range(len(q)):
        rid = server.submit_area(int(q[k]["seed"]), int(q[k]["x"]), int(q[k]["z"]), int(q[k]["half"]))
        r = server.wait(rid)
        assert r.status == 0
        alone_area.append(r.area)
        server.release(rid)
    # the mix: scans interleaved with bursts of area queries, nothing waited on until all are in
    ids = []
    order = list(range(len(q)))
    for i, (st, n) in enumerate(scans):
        ids.append(("s", i, server.submit_scan(st, n)))
        for k in order[i::len(scans)]:
            ids.append(("a", k, server.submit_area(int(q[k]["seed"]), int(q[k]["x"]), int(q[k]["z"]),
                                                   int(q[k]["half"]))))
    for kind, i, rid in ids:
        r = server.wait(rid)
        assert r.status == 0
        if kind == "s":
            assert (server.text(rid), (r.counts.done, r.counts.scanned, r.counts.sites)) == alone_scan[i]
        else:
            a, b = r.area, alone_area[i]
            assert (a.cells, a.sampled, a.minx, a.maxx, a.minz, a.maxz, a.edge, a.border) == \
                (b.cells, b.sampled, b.minx, b.maxx, b.minz, b.maxz, b.edge, b.border)
        assert r.t_submit <= r.t_start <= r.t_done
        server.release(rid)
