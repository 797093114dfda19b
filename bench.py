#!/usr/bin/env python3
"""Benchmark: gated seeds/s of the whole per-seed pipeline on one MI355X, and area-query latency.

Workload (one step): seed indices k in [1e12 + s*1e8, 1e12 + (s+1)*1e8), seeds splitmix64(k), through
the GPU pipeline exactly as `dsearch_desert START COUNT PREFIX THREADS gpu=1 gate=1` runs it: GPU gate
(y-phase <= 0.06f, 4-parity pre-gate >= 0.645, full 512^2 gate >= 0.70) -> octave-0 period tables,
blur and translation scan -> sites -> hot-zone floods -> proxy floods -> exact step-8/step-4 floods ->
CSV rows. Step 0 starts at the 1e7-seed pipeline range from 1e12 and runs ten times its length, so
that the ~780 gated seeds per 1e7 indices keep every pipeline worker busy (steady state). value =
gated seeds (seeds that pass the gate and run through the pipeline: `scanned` of PREFIX.json) per
second over the timed steps. Inputs are seed index ranges generated on the device; outputs are the
CSV files the pipeline writes.

Also reported:
  roofline      the dominant kernel, the persistent tiled flood kernel (k_flood_run, FP64 VALU), per launch
                and (device_level) over the timed region's wall time:
                algorithmic FP64 operations = the reference's sampled cells of the floods it ran x octave
                samples per predicate call of their kind and cost class x 89 FP64 operations per octave
                sample, divided by its device time from HIP events recorded around each launch on the
                stream it runs on (all launches in the timed region). traffic: HBM bytes per launch from
                the rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (calibrated against a
                plain copy), when that summary is present.
  roofline_gate the gate's pre-gate kernel (k_amp<0>, FP32 VALU) in the gate-only leg: 7,077,888 FP32
                operations per phase survivor over its event-timed duration.
  prepass       the same timed steps with the FP32 pre-pass + FP64 confirm (dsx_set_precision 1):
                gated seeds/s, and whether its output rows equal the plain FP64 run's; fp32_error: max
                |FP32 - FP64| of each noise channel at 20,000 random points of 2 seeds.
  cpu_baseline  the reference's own CPU dsearch_desert (built from its sources into oracle/_ref, gpu=0)
                over a bounded sample of seed indices on this host's cores: scanned / seconds. Falls
                back to the oracle's C++ restatement of the same pipeline when oracle/_ref is absent.
                post_gate: the same CPU pipeline in stdin=1 mode on exactly the seeds the GPU gate passes
                over the 1e7-index range [1e12, 1e12 + 1e7), so both sides run the same gated seeds after the gate.
  gate_only     the GPU gate alone over 2^24 seed indices (the gate kernel workload), median of 3 calls,
                with the gate replayed as one hipGraph per chunk (graph_on_s, the default; 2^24 indices here;
                its pre-gate decides from the fused-blur amplitude where that is provably the exact
                decision) and stage by stage in the reference's arithmetic with a survivor-count read-back
                after each stage (graph_off_s), on the same rows (rows_identical); survivors_only_s: the
                graph without amplitudes, as the pipeline calls it (its full gate also decides from a fast
                amplitude where provably exact), and whether its survivors are the same.
  tuning        the autotuner's settings (desert-search_amd/tuned.json, applied by the library) against the
                library defaults on the same timed steps, each in fresh processes, alternating, median of 2:
                gated seeds/s of each and whether their output rows are identical.
  area_queries  4096 exact step-1 area queries (seed splitmix64(j), X, Z uniform in +-3e7, window
                half = floor((sqrt(cap)-1)/2) for caps log-uniform over 1e3..1e7 cells) submitted at
                once to the request server: p50/p99 latency per query and queries/s.
  serving       the request server under Poisson arrivals mixing seed-range scans (70%, length
                floor(10^(3+4u)), random 64-bit start) and area queries (30%): p50/p99 latency per class
                and gated seeds/s.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import statistics
import subprocess
import sys
import tempfile
import threading
import time

REPO = os.path.dirname(os.path.abspath(__file__))
# hardware queues for the 16 pipeline worker streams, the gate and spares (DSX_HW_QUEUES overrides, clamped to
# [1, 28] as desert-search_amd/host/hwq.h does)
os.environ["GPU_MAX_HW_QUEUES"] = str(max(1, min(28, int(os.environ.get("DSX_HW_QUEUES") or 20))))
START = 10 ** 12
N_STEP = 10 ** 8
FP64_PEAK_TFLOPS = 78.6
FLOPS_PER_SAMPLE = 89   # FP64 operations of one octave sample (Perlin::sample + NormalNoise weighting)
KINDS = ("desert_state", "hot", "proxy", "desert")


def pred_samples(kind, hot):
    """Octave samples of one predicate call (kind index as DSX_PRED_*, cost class hot/cold)."""
    if kind == 1:
        return 4
    if kind == 2:
        return 14 if hot else 4
    return 52 if hot else 16


def load_package():
    path = os.path.join(REPO, "desert-search_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("desert_search_amd", path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["desert_search_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def host_cores():
    """CPU cores this process may use: affinity, capped by the cgroup quota (as deploy/prun.sh:6-11)."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            n = min(n, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return n


def _cpu_run(args, pre, stdin_text=None, seeds=None, sample=0):
    """One CPU pipeline run: the reference's dsearch_desert from oracle/_ref when built, else the oracle's
    restatement. Returns (kind, scanned, seconds, passed)."""
    cores = host_cores()
    ref = os.path.join(REPO, "oracle", "_ref", "dsearch_desert")
    if os.access(ref, os.X_OK):
        subprocess.run([ref] + args[:2] + [pre, str(cores)] + args[2:], check=True, input=stdin_text, text=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        m = json.load(open(pre + ".json"))
        return "reference", m["scanned"], m["seconds"], m["passed"]
    import numpy as np
    L = ctypes.CDLL(os.path.join(REPO, "oracle", "liboracle.so"))
    L.or_run.restype = ctypes.c_int
    L.or_run.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                         ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    cnt = (ctypes.c_uint64 * 4)()
    secs = ctypes.c_double()
    if seeds is not None:
        arr = np.ascontiguousarray(seeds, dtype=np.uint64)
        L.or_run(0, 0, arr.ctypes.data, len(arr), 1, cores, pre.encode(), cnt, ctypes.byref(secs), None, 0)
    else:
        L.or_run(0, sample, None, 0, 0, cores, pre.encode(), cnt, ctypes.byref(secs), None, 0)
    return "port", cnt[2], secs.value, cnt[1]


def cpu_baseline(gated_seeds, gpu_value, sample=10 ** 7):
    """Two CPU legs on this host's cores:
    - the reference CPU pipeline with its own CPU gate (gpu=0) over seed indices [0, sample): the whole job
      the reference does without a GPU (value, the reported baseline);
    - post_gate: the reference in stdin=1 mode (dsearch_desert.cpp:571-581) on exactly the seeds the GPU gate
      passes over the range [1e12, 1e12 + 1e7): the same gated seeds the GPU pipeline processes,
      so the ratio compares the post-gate pipeline like for like."""
    cores = host_cores()
    with tempfile.TemporaryDirectory() as td:
        kind, scanned, sec, passed = _cpu_run(["0", str(sample)], os.path.join(td, "cpu"), sample=sample)
        text = "".join(f"{int(x)}\n" for x in gated_seeds)
        kind2, scanned2, sec2, _ = _cpu_run(["0", "0", "stdin=1"], os.path.join(td, "post"), stdin_text=text,
                                            seeds=gated_seeds)
    what = ("the reference's dsearch_desert built from its sources" if kind == "reference"
            else "the oracle's C++ restatement of the reference pipeline")
    value = scanned / sec
    post = scanned2 / sec2
    return {"value": value, "unit": "gated seeds/s", "cores": cores, "kind": kind,
            "sample": f"seed indices [0, {sample:.0e}) through {what} with its CPU gate (gpu=0, production "
                      f"Params): {passed} phase survivors, {scanned} gated seeds through the pipeline in "
                      f"{sec:.1f} s on {cores} threads",
            "gpu_vs_cpu": gpu_value / value,
            "post_gate": {"value": post, "unit": "gated seeds/s", "cores": cores, "kind": kind2,
                          "sample": f"{scanned2} seeds the GPU gate passes over [1e12, 1e12 + 1e7), through the "
                                    f"same CPU pipeline in stdin=1 mode (no CPU gate) in {sec2:.1f} s",
                          "gpu_vs_cpu": gpu_value / post}}


def measured_traffic():
    """HBM bytes per launch of the flood kernel from the committed rocprofv3 --pmc summary, or None."""
    path = os.path.join(REPO, "profiles", "flood_run_traffic.json")
    try:
        return float(json.load(open(path))["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def measured_counters():
    """VALU / FP64 / LDS utilisation of the flood engine and the pre-gate kernel from the committed
    rocprofv3 --pmc summary (kernels run alone under the profiler), or None."""
    path = os.path.join(REPO, "profiles", "kernel_counters.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    keep = {"valu": "valu_busy_pct", "p64": "fp64_frac_of_peak", "p32": "fp32_frac_of_peak", "plds": "lds_frac_of_peak",
            "ldsutil": "lds_util_pct", "bc": "lds_bank_conflicts_per_access", "us": "avg_us_alone"}
    out = {name: {v: d[name][k] for k, v in keep.items() if k in d[name]} for name in ("flood_run", "amp_pregate")
           if name in d}
    out["source"] = d.get("source")
    return out


def percentiles(v):
    import numpy as np
    if len(v) == 0:
        return None, None
    a = np.asarray(v) * 1e3
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


def area_burst(dsx, n=4096, workers=16):
    """The 4096 area queries submitted at once to the request server."""
    q, caps = dsx.area_workload(n)
    s = dsx.Server(0, workers=workers)
    try:
        t0 = time.perf_counter()
        ids = [s.submit_area(int(r["seed"]), int(r["x"]), int(r["z"]), int(r["half"])) for r in q]
        lat, cells = [], 0
        for i in ids:
            r = s.wait(i)
            if r.status != 0:
                raise RuntimeError("area query failed")
            lat.append(r.t_done - r.t_submit)
            cells += r.area.cells
            s.release(i)
        sec = time.perf_counter() - t0
    finally:
        s.close()
    p50, p99 = percentiles(lat)
    return {"queries": n, "seconds": sec, "queries_per_s": n / sec, "p50_ms": p50, "p99_ms": p99,
            "desert_cells": int(cells), "workers": workers}


def serving(dsx, n_req=240, rate=24.0, seed=20261203, workers=16):
    """Poisson arrivals (rate requests/s), 70% scans / 30% area queries, on one request server."""
    import numpy as np
    rng = np.random.default_rng(seed)
    q, caps = dsx.area_workload(n_req)
    gaps = -np.log(1.0 - rng.random(n_req)) / rate
    kinds = rng.random(n_req) < 0.7
    lens = np.floor(10.0 ** (3 + 4 * rng.random(n_req))).astype(np.int64)
    starts = rng.integers(0, 2 ** 63, n_req, dtype=np.int64).astype(np.uint64) * 2 + rng.integers(0, 2, n_req).astype(np.uint64)
    s = dsx.Server(0, workers=workers)
    ids = []
    try:
        t0 = time.perf_counter()
        t = t0
        for k in range(n_req):
            t += gaps[k]
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            if kinds[k]:
                ids.append(("scan", s.submit_scan(int(starts[k]), int(lens[k]))))
            else:
                ids.append(("area", s.submit_area(int(q[k]["seed"]), int(q[k]["x"]), int(q[k]["z"]), int(q[k]["half"]))))
        lat = {"scan": [], "area": []}
        gated = 0
        t_end = t0
        for kind, i in ids:
            r = s.wait(i)
            if r.status != 0:
                raise RuntimeError(f"{kind} request failed")
            lat[kind].append(r.t_done - r.t_submit)
            if kind == "scan":
                gated += r.counts.scanned
            s.release(i)
        sec = time.perf_counter() - t0
    finally:
        s.close()
    out = {"requests": n_req, "rate_per_s": rate, "seconds": sec, "gated_seeds_per_s": gated / sec,
           "workers": workers}
    for kind in ("scan", "area"):
        p50, p99 = percentiles(lat[kind])
        out[kind] = {"n": len(lat[kind]), "p50_ms": p50, "p99_ms": p99}
    return out


def tuning(steps, reps=2):
    """desert-search_amd/tuned.json (the autotuner's settings, applied by dsx_create) against the library
    defaults (DSX_TUNED=0): the workload's timed steps in fresh processes (the settings are read once per
    process), alternating, median of `reps` each; rows_identical: the sorted output rows of every step
    hash the same."""
    path = os.path.join(REPO, "desert-search_amd", "tuned.json")
    env_t = json.load(open(path))["env"] if os.path.exists(path) else {}
    script = os.path.join(REPO, "desert-search_amd", "autotune.py")

    def run(off):
        e = {k: v for k, v in os.environ.items() if k != "DSX_TUNED"}
        if off:
            e["DSX_TUNED"] = "0"
        r = subprocess.run([sys.executable, script, "--trial", "--steps", str(steps)], env=e, capture_output=True,
                           text=True, timeout=300, check=True)
        o = json.loads(r.stdout.strip().splitlines()[-1])
        return o["value"], o["rows"]

    t, d = [], []
    for _ in range(reps):
        t.append(run(False))
        d.append(run(True))
    tv = statistics.median(v for v, _ in t)
    dv = statistics.median(v for v, _ in d)
    return {"env": env_t, "tuned": tv, "default": dv, "speedup": tv / dv, "steps": steps, "reps": reps,
            "rows_identical": len({h for _, h in t + d}) == 1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workers", type=int, default=0, help="pipeline workers (0: the library default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the gate-only, area-query and serving legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device 3 times: the known-size reference for the FETCH_SIZE / "
                         "WRITE_SIZE counters of the HBM traffic profiles")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = 256 << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {n} bytes x 3")
        return
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (seeds shard by index range with no communication; "
              "no multi-GPU run is part of it)", file=sys.stderr)
        sys.exit(2)
    # first, while this process holds no GPU queues of its own (held ones would compete with the trials')
    tun = None if a.no_extra else tuning(a.steps)
    import torch
    pkg = load_package()
    dsx = pkg.dsx
    eng = dsx.Engine(0)
    lib = eng.lib
    params = eng.search_params(gpu=1, gate=1, workers=a.workers)
    td = tempfile.mkdtemp(prefix="dsx_bench_")
    for w in range(a.warmup):
        eng.search(os.path.join(td, f"w{w}"), dsx.SEARCH_GATE, START - (w + 1) * N_STEP, N_STEP, params=params)
    torch.cuda.synchronize()
    lib.dsx_profile(1)
    st = dsx.FloodStats()
    lib.dsx_flood_stats_get(ctypes.byref(st), 1)
    t0 = time.perf_counter()
    gated = 0
    seeds_done = 0
    for s in range(a.steps):
        c = eng.search(os.path.join(td, f"s{s}"), dsx.SEARCH_GATE, START + s * N_STEP, N_STEP, params=params)
        seeds_done += N_STEP
        gated += c[2]
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    lib.dsx_flood_stats_get(ctypes.byref(st), 1)
    lib.dsx_profile(0)
    value = gated / sec
    # roofline of the tile evaluation kernel
    samples = sum(st.sampled_cold[k] * pred_samples(k, False) + st.sampled_hot[k] * pred_samples(k, True)
                  for k in range(4))
    flops = samples * FLOPS_PER_SAMPLE
    launches = max(1, st.eval_launches)
    avg_ms = st.eval_ms / launches
    achieved = flops / (st.eval_ms * 1e-3) / 1e12 if st.eval_ms > 0 else 0.0
    cells_eval = sum(st.evaluated[k] for k in range(4))
    sampled_big = sum(st.sampled_cold[k] + st.sampled_hot[k] for k in range(4))
    out = {
        "metric": "gated seeds/s on 1 MI355X vs host-CPU ref; p50/p99 area-query latency",
        "value": value, "unit": "gated seeds/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * sec / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seed index ranges; seeds = splitmix64(k))",
        "config": {"workload": "whole per-seed pipeline, GPU gate (gate=1) -> tables -> scan -> hot zones -> proxy "
                               "floods -> exact floods; step s = seed indices [1e12 + s*1e8, +1e8)",
                   "seed_indices_per_step": N_STEP, "gated_seeds": int(gated), "workers": int(params.workers) or None},
        "roofline": {"bound": "fp64-valu", "kernel": "k_flood_run (persistent tiled flood engine)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(),
                     "counters": measured_counters(),
                     "avg_launch_ms": avg_ms, "launches": int(st.eval_launches),
                     "flops_per_launch": flops / launches,
                     "octave_samples": int(samples), "sampled_cells": int(sampled_big),
                     "cells_evaluated": int(cells_eval),
                     # the same algorithmic FP64 operations over the timed region's wall time: the engine's
                     # rate at device level, with the launches of the 16 worker streams overlapping
                     "device_level": {"achieved": flops / sec / 1e12, "frac": flops / sec / 1e12 / FP64_PEAK_TFLOPS,
                                      "launch_overlap": st.eval_ms * 1e-3 / sec},
                     "by_kind": {KINDS[k]: {"floods": st.floods[k], "tiles": st.tiles[k], "evaluated": st.evaluated[k],
                                            "sampled_cold": st.sampled_cold[k], "sampled_hot": st.sampled_hot[k],
                                            "by_size": [[st.size_floods[k][b], st.size_tiles[k][b],
                                                         st.size_sampled[k][b]] for b in range(6)]}
                                 for k in range(4) if st.floods[k]},
                     "small_window_floods": st.small_floods, "small_window_sampled": st.small_sampled,
                     "eval_rounds": st.eval_rounds,
                     "needed_per_round": st.eval_needed / max(1, st.eval_rounds),
                     # where the persistent kernel's workgroups spend their time (fractions of their lifetime)
                     "workgroup_time": {"seconds": st.wg_life * 1e-8,
                                        "waiting_for_items": st.wg_wait / max(1, st.wg_life),
                                        "evaluating": st.wg_eval / max(1, st.wg_life),
                                        "loading_octaves": st.wg_load / max(1, st.wg_life)}},
    }
    # the FP32 pre-pass variant over the same steps: throughput and identical rows
    pp = eng.search_params(gpu=1, gate=1, workers=a.workers, prepass=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gated_pp = 0
    for s in range(a.steps):
        gated_pp += eng.search(os.path.join(td, f"p{s}"), dsx.SEARCH_GATE, START + s * N_STEP, N_STEP, params=pp)[2]
    torch.cuda.synchronize()
    sec_pp = time.perf_counter() - t0
    same = all(sorted(open(os.path.join(td, f"{t}{s}.csv")).read().splitlines()) ==
               sorted(open(os.path.join(td, f"p{s}.csv")).read().splitlines()) and
               sorted(open(os.path.join(td, f"{t}{s}.seeds.csv")).read().splitlines()) ==
               sorted(open(os.path.join(td, f"p{s}.seeds.csv")).read().splitlines())
               for s in range(a.steps) for t in ["s"])
    import numpy as np
    rng = np.random.default_rng(3)
    err = {}
    for seed in (0x5eed, 3008578943537042204):
        pts = np.zeros((20000, 3))
        pts[:, 0] = rng.integers(-7_500_000, 7_500_000, len(pts))
        pts[:, 2] = rng.integers(-7_500_000, 7_500_000, len(pts))
        for ch, name in enumerate(("T", "H", "C", "E", "shift", "W")):
            d = float(np.abs(eng.noise_f32(seed, ch, pts) - eng.noise(seed, ch, pts)).max())
            err[name] = max(err.get(name, 0.0), d)
    out["prepass"] = {"value": gated_pp / sec_pp, "ms_per_step": 1e3 * sec_pp / a.steps, "gated_seeds": int(gated_pp),
                      "outputs_identical": bool(same and gated_pp == gated), "vs_fp64": (gated_pp / sec_pp) / value,
                      "fp32_error": err, "bound_T": 2e-4}
    # the seeds the GPU gate passes over [1e12, 1e12 + 1e7): the CPU baseline's post-gate leg runs exactly these
    cfg3_gated = eng.gate(START, 10 ** 7)[0] if not a.no_cpu else None
    eng.close()
    if not a.no_extra:
        g = dsx.Engine(0)
        gsec, rows = {}, {}
        for mode in (1, 0, 1, 0):   # graph replay, then stage by stage, twice each; the same 2^24 rows
            g.set_gate_graph(mode)
            g.gate(0, 1 << 24)
            reps = []
            for _ in range(3):
                t1 = time.perf_counter()
                sd, am, _, cnt = g.gate(0, 1 << 24)
                reps.append(time.perf_counter() - t1)
            gsec[mode] = min(gsec.get(mode, 1e9), float(np.median(reps)))
            rows[mode] = (sd.tobytes(), am.tobytes(), tuple(cnt))
        ms, work = g.gate_profile()   # per-kernel times of the last (stage-by-stage) call
        g.set_gate_graph(1)
        # survivors only, as the pipeline calls it (no amplitudes: the full gate decides from its fast
        # amplitude too); the same survivors
        g.gate(0, 1 << 24, want_amps=False)
        reps = []
        for _ in range(3):
            t1 = time.perf_counter()
            so = g.gate(0, 1 << 24, want_amps=False)
            reps.append(time.perf_counter() - t1)
        so_s = float(np.median(reps))
        g.close()
        out["gate_only"] = {"seed_indices": 1 << 24, "seconds": gsec[1], "gated_seeds_per_s": cnt[2] / gsec[1],
                            "counts": list(cnt), "graph_on_s": gsec[1], "graph_off_s": gsec[0],
                            "graph_speedup": gsec[0] / gsec[1], "rows_identical": rows[0] == rows[1],
                            "survivors_only_s": so_s, "survivors_only_gated_seeds_per_s": so[3][2] / so_s,
                            "survivors_only_identical": so[0].tobytes() == rows[1][0] and tuple(so[3]) == rows[1][2]}
        gach = 7077888 * work[2] / (ms[2] * 1e-3) / 1e12
        out["roofline_gate"] = {"bound": "fp32-valu", "kernel": "k_amp<0> (4-parity pre-gate)", "achieved": gach,
                                "peak": 157.3, "unit": "TFLOP/s", "frac": gach / 157.3, "avg_launch_ms": ms[2],
                                "phase_survivors": int(work[2])}
        out["tuning"] = tun
        out["area_queries"] = area_burst(dsx)
        out["serving"] = serving(dsx)
    # flood batches that hit the engine's time limit (run again on a reset engine) or exhausted its tile pool,
    # over every leg of this run: 0 in a healthy run
    st2 = dsx.FloodStats()
    lib.dsx_flood_stats_get(ctypes.byref(st2), 1)
    out["flood_health"] = {"time_retries": int(st.time_retries + st2.time_retries),
                           "capacity_resets": int(st.capacity_resets + st2.capacity_resets)}
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline(cfg3_gated, value)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
