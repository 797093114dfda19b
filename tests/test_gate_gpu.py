"""GPU seed gate (libdsx.so through the C ABI) against the CPU oracle: bit-exact survivor sets and
amplitudes; batching invariance and the published pre-gate recall at the full 2^24 size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("start", [0, 10 ** 12, (1 << 64) - (1 << 19)])
def test_gate_matches_oracle(engine, oracle, start):
    """Survivor set, amplitudes (bitwise) and per-stage counts over 2^19 indices (wrapping u64 too)."""
    count = 1 << 19
    seeds, amps, k, cnt = engine.gate(start, count)
    ok, oamps, ocnt = oracle.gate_run(start, count)
    assert cnt == ocnt
    assert np.array_equal(k, ok)
    assert np.array_equal(amps.view(np.uint32), oamps.view(np.uint32))
    assert all(int(s) == oracle.or_mix((start + int(i)) % (1 << 64)) for s, i in zip(seeds, k))


def test_gate_amplitudes_bitexact(engine, oracle):
    """Pre-gate (4-parity and legacy) and full amplitudes of arbitrary seeds, bitwise."""
    pkg_params = engine.lib  # noqa: F841
    rng = np.random.default_rng(7)
    seeds = rng.integers(0, 1 << 63, size=96, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    pre, full = engine.gate_amps(seeds)
    for s, p, f in zip(seeds, pre, full):
        assert np.float32(oracle.or_gate_amp(int(s), 256, 4, 1, 1.0)).view(np.uint32) == p.view(np.uint32)
        assert np.float32(oracle.or_gate_amp(int(s), 512, 7, 0, 2.0)).view(np.uint32) == f.view(np.uint32)
    from conftest import load_package
    dsx = load_package().dsx
    pre0, _ = engine.gate_amps(seeds[:32], dsx.GateParams(box=0))
    for s, p in zip(seeds[:32], pre0):
        assert np.float32(oracle.or_gate_amp(int(s), 256, 4, 0, 1.0)).view(np.uint32) == p.view(np.uint32)


def test_gate_legacy_mode_matches_oracle(engine, oracle):
    """box=0: even-parity pre-gate at amp - margin (gate.cuh:392), then the full gate."""
    from conftest import load_package
    dsx = load_package().dsx
    start, count = 77, 1 << 18
    seeds, amps, k, cnt = engine.gate(start, count, dsx.GateParams(box=0))
    ks = oracle.phase_list(start, count, 0.06)
    want = []
    thr = np.float32(np.float32(0.70) - np.float32(0.01))
    for kk in ks:
        s = oracle.or_mix(start + int(kk))
        if oracle.or_gate_amp(s, 256, 4, 0, 1.0) >= thr:
            a = oracle.or_gate_amp(s, 512, 7, 0, 2.0)
            if a >= np.float32(0.70):
                want.append(int(kk))
    assert [int(x) for x in k] == want
    assert cnt[0] == len(ks)


def test_gate_full_size_batch_invariance(engine):
    """The 2^24-index gate workload: one call equals 16 calls of 2^20 (a checksum of checksums),
    survivors sorted and above threshold."""
    seeds, amps, k, cnt = engine.gate(0, 1 << 24)
    assert np.all(np.diff(k.astype(np.int64)) > 0)
    assert np.all(amps >= np.float32(0.70))
    parts = [engine.gate(i << 20, 1 << 20) for i in range(16)]
    k2 = np.concatenate([p[2] + np.uint64(i << 20) for i, p in enumerate(parts)])
    a2 = np.concatenate([p[1] for p in parts])
    assert np.array_equal(k, k2) and np.array_equal(amps.view(np.uint32), a2.view(np.uint32))
    assert cnt == tuple(sum(p[3][j] for p in parts) for j in range(3))
    # ~0.72% phase survivors, ~1 gated seed per 13,000 indices (README.md:97)
    assert abs(cnt[0] / (1 << 24) - 0.0072) < 0.0003
    assert 900 < cnt[2] < 1900


def test_pregate_recall(engine, oracle):
    """Recall of the 4-parity pre-gate at pthr 0.645 for seeds whose full amplitude is >= 0.70:
    published 99.5% (PROGRESSION.md:107-111)."""
    ks = oracle.phase_list(0, 1 << 22, 0.06)
    seeds = np.array([oracle.or_mix(int(k)) for k in ks], np.uint64)
    pre, full = engine.gate_amps(seeds)
    hot = full >= np.float32(0.70)
    assert hot.sum() > 200
    assert (pre[hot] >= np.float32(0.645)).mean() >= 0.97


def test_gate_capacity_and_argument_errors(engine):
    from conftest import load_package
    dsx = load_package().dsx
    seeds, amps, k, cnt = engine.gate(0, 1 << 20)
    assert len(k) > 2
    with pytest.raises(dsx.DsxError) as e:
        engine.gate(0, 1 << 20, cap=2)
    assert e.value.code == dsx.DSX_E_CAPACITY
    with pytest.raises(dsx.DsxError) as e:
        engine.gate(0, 10, dsx.GateParams(box=5))
    assert e.value.code == dsx.DSX_E_ARG
    s, a, kk, c = engine.gate(5, 0)
    assert len(kk) == 0 and c == (0, 0, 0)


def test_tgate_tool(engine, oracle):
    """The tgate command line (tgate.cu:1-7): "seed amp" lines equal to dsx_gate's survivors, check=1
    prints the GPU amplitude beside tgate.cu:14-39's CPU amplitude for the first 4 (different arithmetic:
    equal to a few 1e-6), calib=N gives "seed pre full phase" of the first N phase survivors (by seed
    value) of 2^26 indices."""
    import ctypes
    import os
    import subprocess
    from conftest import REPO
    exe = os.path.join(REPO, "desert-search_amd", "tgate")
    start, count = 10 ** 12, 1 << 20
    r = subprocess.run([exe, str(start), str(count), "check=1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    from conftest import load_package
    dsx = load_package().dsx
    seeds, amps, k, cnt = engine.gate(start, count, dsx.GateParams(max_phase=0.10))   # GateCfg::maxph default
    assert r.stdout.splitlines() == ["%d %.5f" % (int(s), float(a)) for s, a in zip(seeds, amps)]
    checks = [l for l in r.stderr.splitlines() if "check seed" in l]
    assert len(checks) == min(4, len(seeds))
    for l in checks:
        w = l.split()   # "check seed S: gpu amp A cpu amp B"
        assert w[3:5] == ["gpu", "amp"] and w[6:8] == ["cpu", "amp"]
        assert abs(float(w[5]) - float(w[8])) <= 2e-5, l
    assert "%d seed indices, %d phase, %d pre-gate, %d gated" % ((count,) + tuple(cnt)) in r.stderr
    # amps=0: the seeds alone (the library is asked for no amplitudes), the same seeds
    r0 = subprocess.run([exe, str(start), str(count), "amps=0"], capture_output=True, text=True, timeout=120)
    assert r0.returncode == 0, r0.stderr
    assert r0.stdout.splitlines() == ["%d" % int(s) for s in seeds]
    n = 12
    r = subprocess.run([exe, "5", "1", "calib=%d" % n, "phase=0.10"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ks = oracle.phase_list(5, 1 << 26, 0.10)
    want = sorted(oracle.or_mix(5 + int(x)) for x in ks)[:n]
    lines = r.stdout.splitlines()
    assert [int(l.split()[0]) for l in lines] == want
    for l in lines:
        s, p, f, ph = l.split()
        assert p == "%.5f" % oracle.or_gate_amp(int(s), 256, 4, 1, 1.0)
        assert f == "%.5f" % oracle.or_gate_amp(int(s), 512, 7, 0, 2.0)
        fy = (ctypes.c_double * 2)()
        assert ph == "%.5f" % oracle.or_phase(int(s), ctypes.byref(fy))
    r = subprocess.run([exe, "0", "10", "bogus=1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown key bogus" in r.stderr


def test_gate_graph_equals_eager(engine):
    """The gate replayed as a hipGraph per chunk (dsx_set_gate_graph(1), the default) gives the survivors,
    amplitudes and counts of the stage-by-stage launches, bit for bit: across chunk boundaries and a partial
    last chunk (2^24 + 777 at max_phase 0.06: chunks of 2^24), the 64-bit index wrap, calls of other sizes
    and phase cuts (graph rebuilt; 0.3 gives 2^20-index chunks), a phase cut whose survivors overflow the
    graph's grid (max_phase 0.9: about 98 % survive, the grid holds 0.625 of a chunk + 4096, so every chunk
    is redone stage by stage), and the legacy pre-gate; graph replays alternate with eager calls on the same
    context."""
    from conftest import load_package
    dsx = load_package().dsx
    cases = [(10 ** 12, (1 << 24) + 777, dsx.GateParams()), (0, 1 << 22, dsx.GateParams()),
             ((1 << 64) - (1 << 20), 1 << 21, dsx.GateParams()), (5 * 10 ** 11, 3 << 20, dsx.GateParams(max_phase=0.10)),
             (7, 1 << 21, dsx.GateParams(max_phase=0.3)), (10 ** 12, 1 << 22, dsx.GateParams(box=0)),
             (3, 1 << 20, dsx.GateParams(max_phase=0.9))]
    try:
        for rep in range(2):
            for start, count, p in cases:
                engine.set_gate_graph(1)
                g = engine.gate(start % (1 << 64), count, p)
                engine.set_gate_graph(0)
                e = engine.gate(start % (1 << 64), count, p)
                engine.set_gate_graph(1)
                f = engine.gate(start % (1 << 64), count, p, want_amps=False)
                print(f"gate graph vs eager: start={start} count={count} counts {g[3]} / {e[3]} / survivors only {f[3]}",
                      flush=True)
                assert g[3] == e[3] == f[3]
                assert np.array_equal(g[0], e[0]) and np.array_equal(g[2], e[2])
                assert np.array_equal(g[1].view(np.uint32), e[1].view(np.uint32))
                assert np.array_equal(f[0], e[0]) and np.array_equal(f[2], e[2]) and f[1] is None
                assert len(g[0]) > 0
                if p.max_phase > 0.5:
                    assert g[3][0] > 0.7 * count   # beyond the graph grid's 0.625 of a chunk + 4096
    finally:
        engine.set_gate_graph(1)


def test_fast_amplitude_within_margin(engine):
    """Each fast pre-gate amplitude (1: FMA blurs; 2: FMA blurs of radius 3; 3: radius 3 over the grid summed
    in 16-bit fixed point, blurred in integer dot products), and the full gate's fast amplitude (the 512^2
    grid in 16 bits, integer passes at radius 7), lies within its decision margin of the reference amplitude,
    on 2^18 random seeds (any seed has both amplitudes); the fast kernels only decide seeds farther than that
    margin from the threshold, so this is what keeps their decisions the reference's. The margins grow with
    each approximation and the fixed-point amplitudes differ from the float ones."""
    import os
    seeds = np.random.default_rng(11).integers(0, 1 << 63, 1 << 18, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    exact, _ = engine.gate_amps(seeds)
    fast, eps = {}, {}
    try:
        for dec in (1, 2, 3):
            os.environ["DSX_GATE_AMPS_FAST"] = str(dec)
            fast[dec], eps[dec] = engine.gate_amps(seeds)
    finally:
        os.environ.pop("DSX_GATE_AMPS_FAST", None)
    for dec in (1, 2, 3):
        d = np.abs(fast[dec].astype(np.float64) - exact.astype(np.float64))
        r = d / eps[dec].astype(np.float64)
        print(f"fast amplitude {dec}: max |fast - exact| {d.max():.3e}, max ratio to margin {r.max():.3f}, "
              f"median margin {np.median(eps[dec]):.3e}", flush=True)
        assert np.all(np.isfinite(fast[dec])) and np.all(eps[dec] > 0)
        assert np.all(d <= eps[dec])
    assert np.all(eps[2] >= eps[1]) and np.all(eps[3] > eps[2])
    assert not np.array_equal(fast[2].view(np.uint32), fast[3].view(np.uint32))
    # the full gate's fast amplitude (16-bit grid, integer passes at radius 7) on the same seeds
    _, exact_full = engine.gate_amps(seeds)
    try:
        os.environ["DSX_GATE_AMPS_FAST_FULL"] = "1"
        eps_full, fast_full = engine.gate_amps(seeds)
    finally:
        os.environ.pop("DSX_GATE_AMPS_FAST_FULL", None)
    d = np.abs(fast_full.astype(np.float64) - exact_full.astype(np.float64))
    print(f"fast full amplitude: max |fast - exact| {d.max():.3e}, max ratio to margin "
          f"{(d / eps_full.astype(np.float64)).max():.3f}, median margin {np.median(eps_full):.3e}", flush=True)
    assert np.all(np.isfinite(fast_full)) and np.all(eps_full > 0)
    assert np.all(d <= eps_full)
    assert not np.array_equal(fast_full.view(np.uint32), exact_full.view(np.uint32))


def test_fast_decisions_at_exact_thresholds(engine):
    """The graph chain decides a seed with the fast amplitude (FMA blurs) only when that amplitude is
    farther from the threshold than its proven error bound, and runs the reference's arithmetic for the
    rest and for every amplitude it outputs. Thresholds set exactly to a surviving seed's pre-gate or full
    amplitude (amp - thr == 0, which the fast amplitude can never decide) keep that seed; the next float
    above removes it. Each case: the graph chain equals the stage-by-stage exact launches bit for bit, and
    the boundary seed is in or out as its explicit amplitudes (dsx_gate_amps) say. Pre-gate boundaries in
    the anti-aliased (box=1, pthr) and legacy (box=0, amp - margin: a Sterbenz-exact margin) modes; full
    gate boundaries through amp (box=1)."""
    from conftest import load_package
    dsx = load_package().dsx
    start, count = 10 ** 12, 1 << 21
    up = lambda x: float(np.nextafter(np.float32(x), np.float32(np.inf)))
    try:
        for box in (1, 0):
            base = dsx.GateParams(box=box)
            engine.set_gate_graph(0)
            seeds, amps, _, _ = engine.gate(start, count, base)
            pre, full = engine.gate_amps(seeds, base)
            assert np.array_equal(full.view(np.uint32), amps.view(np.uint32))
            i, j = int(np.argmin(pre)), int(np.argmin(full))
            cases = []
            for t, keep in ((float(pre[i]), True), (up(pre[i]), False)):
                if box:
                    cases.append((dsx.GateParams(box=1, pthr=t), seeds[i], keep))
                else:
                    m = np.float32(base.amp) - np.float32(t)   # exact (Sterbenz): amp - margin == t in float
                    assert np.float32(np.float32(base.amp) - m) == np.float32(t)
                    cases.append((dsx.GateParams(box=0, margin=float(m)), seeds[i], keep))
            if box:
                cases += [(dsx.GateParams(box=1, amp=float(full[j])), seeds[j], True),
                          (dsx.GateParams(box=1, amp=up(full[j])), seeds[j], False)]
            for p, s, keep in cases:
                engine.set_gate_graph(1)
                g = engine.gate(start, count, p)
                engine.set_gate_graph(0)
                e = engine.gate(start, count, p)
                engine.set_gate_graph(1)
                f = engine.gate(start, count, p, want_amps=False)   # survivors only: the full gate decided fast
                print(f"boundary box={box} pthr={p.pthr!r} margin={p.margin!r} amp={p.amp!r}: counts {g[3]} / {e[3]}"
                      f" / {f[3]}", flush=True)
                assert g[3] == e[3] == f[3]
                assert np.array_equal(g[0], e[0]) and np.array_equal(g[1].view(np.uint32), e[1].view(np.uint32))
                assert np.array_equal(f[0], e[0]) and np.array_equal(f[2], e[2])
                assert (int(s) in set(g[0].tolist())) == keep
    finally:
        engine.set_gate_graph(1)
