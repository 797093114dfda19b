"""Autotuner of libdsx's launch and batching settings on the whole per-seed pipeline.

    python desert-search_amd/autotune.py [--steps 4] [--min-gain 0.01] [--passes 2] [--confirm 3] [--dry-run]
                                         [--out PATH] [--knobs NAME,...] [--from-tuned]

Workload: the search in gate mode (gate=1: GPU gate -> tables -> scan -> hot zones -> proxy floods ->
exact floods) over seed indices [1e12 + s*1e8, +1e8) for s < steps, after one warm-up step on the range
before them; score = gated seeds per second over the timed steps. Every setting swept here is
result-invariant (grid sizes and workgroup shapes of the persistent flood engine, pipeline workers, host
wait mode, seed chunk and group sizes, flood routing): each trial's output rows (the sorted CSV lines of
every timed step) must hash to the default run's, and a setting whose rows differ is rejected and
reported.

The settings are environment variables read once per process, so each trial is a fresh process. Search:
three runs of the starting point (the library defaults, or with --from-tuned the settings tuned.json holds
now) give the run-to-run spread; then coordinate descent, each knob in turn over its candidates with the
others at the best so far, --passes times or until a pass changes nothing; a change is kept when it is
faster than the best so far by more than the threshold max(--min-gain, spread). A knob named "A+B" moves
the variables A and B together (candidate "x+y" sets A=x, B=y): settings that only pay together, such as
more pipeline workers once the host threads block instead of spinning while they wait for the device, are
out of reach of one-variable moves. --knobs restricts the sweep to the named knobs. The winner and the
library defaults are then timed in alternation (--confirm pairs), and tuned.json (beside libdsx.so, read
by dsx_create; an environment variable that is set wins, DSX_TUNED=0 ignores the file) holds the winner
only when its median beats the defaults' median by the threshold, else an empty "env".
"""
import argparse
import hashlib
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "tuned.json")
START = 10 ** 12
N_STEP = 10 ** 8

# (variable, candidates); "" = unset, the library default. Orders: the knobs with the largest effect first.
KNOBS = [
    ("DSX_WORKERS", ["", "12", "20", "24"]),                    # pipeline workers (default 16)
    ("DSX_SYNC", ["", "spin", "yield", "block"]),               # host wait mode of stream synchronisations
    ("DSX_SYNC+DSX_WORKERS", ["block+20", "block+24", "block+28"]),   # blocking waits free host cores for workers
    ("DSX_FLOOD_NT", ["", "128", "512"]),                       # light flood instance workgroup (256)
    ("DSX_FLOOD_NT_FULL", ["", "256", "128"]),                  # full (exact) instance workgroup (512)
    ("DSX_FLOOD_OCC", ["", "4", "5", "6"]),                     # occupancy hint of the 256-thread instances
    ("DSX_FLOOD_BLOCKS_PER_JOB", ["", "2", "8"]),               # light launch width per tiled flood (4)
    ("DSX_FLOOD_MAX_BLOCKS", ["", "64", "256"]),                # light launch cap (128)
    ("DSX_FLOOD_BLOCKS_PER_EXACT", ["", "2", "8"]),             # full launch width per exact flood (4)
    ("DSX_FLOOD_MIN_BLOCKS_FULL", ["", "8", "32"]),             # full launch floor (16)
    ("DSX_FLOOD_MAX_BLOCKS_FULL", ["", "64", "256"]),           # full launch cap (128)
    ("DSX_FLOOD_TILES_PER_UNIT", ["", "4", "16"]),              # hinted tiles per extra 256-thread unit (8)
    ("DSX_LDS_PROXY_SMALL", ["", "1"]),                         # proxy floods on the LDS engine first (0)
    ("DSX_LDS_HOT_SMALL", ["", "0"]),                           # hot-zone floods on the LDS engine first (1)
    ("DSX_GROUP", ["", "32", "128"]),                           # seeds per worker group (64)
    ("DSX_ACTIVE", ["", "2"]),                                  # continuous batching: seeds in flight (1)
    ("DSX_CHUNK", ["", "2097152", "8388608"]),                  # seed indices per gate call (2^22)
    ("DSX_FLOOD_SPEC", ["", "0"]),                              # speculative sub-blocks in idle waves (1)
    ("DSX_GATE_TRUNC", ["", "0"]),                              # truncated-blur fast pre-gate decisions (1)
]


def knob_vars(k):
    return k.split("+")


def knob_vals(k, v):
    """The values candidate v of knob k gives its variables ("" = all unset)."""
    n = len(knob_vars(k))
    vals = v.split("+") if v != "" else [""] * n
    if len(vals) != n:
        raise ValueError(f"candidate {v!r} of knob {k} sets {len(vals)} values for {n} variables")
    return dict(zip(knob_vars(k), vals))


def load_dsx():
    import importlib.util
    path = os.path.join(HERE, "__init__.py")
    spec = importlib.util.spec_from_file_location("desert_search_amd", path, submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["desert_search_amd"] = mod
    spec.loader.exec_module(mod)
    return mod.dsx


def run_steps(dsx, steps, td):
    """One warm-up step, then `steps` timed steps of the workload; returns (gated seeds/s, rows hash)."""
    eng = dsx.Engine(0)
    try:
        params = eng.search_params(gpu=1, gate=1)
        eng.search(os.path.join(td, "w"), dsx.SEARCH_GATE, START - N_STEP, N_STEP, params=params)
        gated = 0
        t0 = time.perf_counter()
        for s in range(steps):
            gated += eng.search(os.path.join(td, f"s{s}"), dsx.SEARCH_GATE, START + s * N_STEP, N_STEP,
                                params=params)[2]
        sec = time.perf_counter() - t0
    finally:
        eng.close()
    h = hashlib.sha256()
    for s in range(steps):
        with open(os.path.join(td, f"s{s}.csv")) as f:
            h.update("".join(sorted(f.readlines())).encode())
    return gated / sec, h.hexdigest()


def trial(env, steps):
    """One trial in a fresh process with `env` (variables "" are removed); returns (value, rows hash)."""
    e = dict(os.environ)
    e["DSX_TUNED"] = "0"
    for k, v in env.items():
        if v == "":
            e.pop(k, None)
        else:
            e[k] = v
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--trial", "--steps", str(steps)], env=e,
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"trial {env} failed ({r.returncode}): {r.stderr[-2000:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    return out["value"], out["rows"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--min-gain", type=float, default=0.01)
    ap.add_argument("--passes", type=int, default=2, help="coordinate-descent passes (stops early when one changes nothing)")
    ap.add_argument("--confirm", type=int, default=3)
    ap.add_argument("--dry-run", action="store_true", help="search, but do not write tuned.json")
    ap.add_argument("--out", default=TUNED, help="where to write the result (default: tuned.json beside libdsx.so)")
    ap.add_argument("--knobs", default="", help="comma-separated knob names to sweep (default: all)")
    ap.add_argument("--from-tuned", action="store_true", help="start the descent from the settings tuned.json holds")
    ap.add_argument("--trial", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.trial:
        with tempfile.TemporaryDirectory(prefix="dsx_tune_") as td:
            v, rows = run_steps(load_dsx(), a.steps, td)
        print(json.dumps({"value": v, "rows": rows}))
        return
    base = {x: "" for k, _ in KNOBS for x in knob_vars(k)}
    knobs = KNOBS
    if a.knobs:
        names = a.knobs.split(",")
        unknown = set(names) - {k for k, _ in KNOBS}
        if unknown:
            raise SystemExit(f"unknown knobs: {sorted(unknown)}")
        knobs = [(k, c) for k, c in KNOBS if k in names]
    start = dict(base)
    if a.from_tuned:
        with open(TUNED) as f:
            start.update(json.load(f).get("env", {}))
    # run-to-run spread of the starting point: a change must beat the best so far by more than it
    reps = [trial(start, a.steps) for _ in range(3)]
    rows0 = reps[0][1]
    if any(r[1] != rows0 for r in reps):
        raise RuntimeError("the starting point's runs' rows differ from each other")
    vals = sorted(r[0] for r in reps)
    best_v = statistics.median(vals)
    spread = (vals[-1] - vals[0]) / best_v
    gain = max(a.min_gain, spread)
    print(f"start: {best_v:.1f} gated seeds/s (spread {100 * spread:.1f} %, threshold {100 * gain:.1f} %)",
          flush=True)
    best = dict(start)
    log, rejected = [{"env": {x: y for x, y in start.items() if y != ""}, "value": v} for v in vals], []
    for p in range(a.passes):
        changed_pass = False
        for k, cands in knobs:
            for v in cands:
                kv = knob_vals(k, v)
                if all(best[x] == y for x, y in kv.items()):
                    continue
                env = dict(best, **kv)
                val, rows = trial(env, a.steps)
                changed = {x: y for x, y in env.items() if y != ""}
                log.append({"env": changed, "value": val})
                if rows != rows0:
                    rejected.append(changed)
                    print(f"  {k}={v}: {val:.1f}  ROWS DIFFER - rejected", flush=True)
                    continue
                keep = val > best_v * (1 + gain)
                print(f"  pass {p}: {k}={v}: {val:.1f}{'  kept' if keep else ''}", flush=True)
                if keep:
                    best, best_v, changed_pass = env, val, True
        if not changed_pass:
            break
    tuned = {x: y for x, y in best.items() if y != ""}
    # confirmation: the defaults and the winner in alternation
    dv, tv = [], []
    if tuned:
        for _ in range(a.confirm):
            dv.append(trial(base, a.steps)[0])
            tv.append(trial(best, a.steps)[0])
        d_med, t_med = statistics.median(dv), statistics.median(tv)
        print(f"confirm: default {d_med:.1f}, tuned {t_med:.1f} ({t_med / d_med:.3f}x)", flush=True)
        if t_med < d_med * (1 + gain):
            tuned = {}
    doc = {"env": tuned,
           "device": "MI355X (gfx950), one GPU",
           "start": {x: y for x, y in start.items() if y != ""}, "knobs": [k for k, _ in knobs],
           "workload": f"search gate=1 over seed indices [1e12 + s*1e8, +1e8), s < {a.steps}, after 1 warm-up step",
           "unit": "gated seeds/s", "min_gain": gain, "start_spread": spread,
           "default_value": statistics.median(dv) if dv else best_v if not tuned else None,
           "tuned_value": statistics.median(tv) if tv else None,
           "confirm": {"default": dv, "tuned": tv}, "rejected_rows_differ": rejected, "trials": log}
    print(json.dumps({"env": tuned, "default": doc["default_value"], "tuned": doc["tuned_value"]}), flush=True)
    if not a.dry_run:
        with open(a.out, "w") as f:
            json.dump(doc, f, indent=1)
            f.write("\n")
        print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    main()
