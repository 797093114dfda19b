#!/usr/bin/env python3
"""Benchmark: gated seeds/s of the GPU seed gate on one MI355X.

Workload (one step): seed indices k in [0, 2^24), seeds splitmix64(k), through the GPU gate - y-phase
<= 0.06f, 4-parity anti-aliased 256^2 pre-gate >= 0.645, full 512^2 amplitude gate >= 0.70 - with
the survivor list sorted and copied back. value = gated seeds (full-gate survivors) per second.
Inputs are generated on the device (seed index ranges), so nothing crosses PCIe before a step.

Also reported:
  roofline      the dominant kernel (the pre-gate k_amp<0>): algorithmic FP32 operations per launch
                (18 adds of the value grid + 2 x 9 multiply-adds of the blur per cell, 2 x 65,536
                cells per phase survivor) / its average launch time, from HIP events recorded by the
                library on the stream the kernel runs on, against MI355X's published FP32 vector peak.
  cpu_baseline  the same gate computed by the CPU oracle (a C++ port of the identical arithmetic) on a
                bounded sample of the same workload, on this host's cores.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
FLOPS_PER_PRE_SEED = 2 * 65536 * (18 + 2 * 9 * 2)   # 7,077,888 FP32 ops per phase survivor
FP32_PEAK_TFLOPS = 157.3
N_SEEDS = 1 << 24


def load_package():
    path = os.path.join(REPO, "desert-search_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("desert_search_amd", path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["desert_search_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(sample=1 << 21):
    """Oracle GPU-gate restatement over seed indices [0, sample) on this host's cores."""
    L = ctypes.CDLL(os.path.join(REPO, "oracle", "liboracle.so"))
    L.or_gate_run.restype = ctypes.c_int64
    L.or_gate_run.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    cores = min(16, len(os.sched_getaffinity(0)))
    cap = 1 << 14
    idx = (ctypes.c_uint64 * cap)()
    amp = (ctypes.c_float * cap)()
    cnt = (ctypes.c_uint64 * 3)()
    t = time.perf_counter()
    L.or_gate_run(0, sample, 0.06, 0.645, 0.70, cores, idx, amp, cap, cnt)
    sec = time.perf_counter() - t
    return {"value": cnt[2] / sec, "unit": "gated seeds/s", "cores": cores, "kind": "port",
            "sample": f"seed indices [0, 2^{sample.bit_length() - 1}) through the oracle's C++ restatement of the "
                      f"same gate ({cnt[0]} phase, {cnt[1]} pre-gate, {cnt[2]} gated seeds in {sec:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the seed gate shards by seed range; no multi-GPU "
              "run is part of it)", file=sys.stderr)
        sys.exit(2)
    import torch
    pkg = load_package()
    eng = pkg.dsx.Engine(0)
    for _ in range(a.warmup):
        eng.gate(0, N_SEEDS)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gated = 0
    pre_ms = 0.0
    pre_work = 0
    for _ in range(a.steps):
        seeds, amps, k, cnt = eng.gate(0, N_SEEDS)
        gated += cnt[2]
        ms, work = eng.gate_profile()
        pre_ms += ms[2]
        pre_work += int(work[2])
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    value = gated / sec
    avg_launch_ms = pre_ms / a.steps
    achieved = FLOPS_PER_PRE_SEED * (pre_work / a.steps) / (avg_launch_ms * 1e-3) / 1e12
    out = {
        "metric": "gated seeds/s on 1 MI355X vs host-CPU ref; p50/p99 area-query latency",
        "value": value, "unit": "gated seeds/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * sec / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seed indices; seeds = splitmix64(k))",
        "config": {"workload": "seed gate, BASELINE config 2: k in [0, 2^24), y-phase <= 0.06f, 4-parity "
                               "pre-gate >= 0.645, full 512^2 gate >= 0.70",
                   "seeds_per_step": N_SEEDS, "gated_per_step": cnt[2], "phase_survivors_per_step": cnt[0],
                   "pregate_survivors_per_step": cnt[1]},
        "roofline": {"bound": "valu", "kernel": "k_amp<0> (4-parity pre-gate)", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": None, "avg_launch_ms": avg_launch_ms,
                     "flops_per_launch": FLOPS_PER_PRE_SEED * (pre_work / a.steps)},
        "cpu_baseline": None if a.no_cpu else cpu_baseline(),
    }
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
