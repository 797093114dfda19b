#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ by running the reference's own CPU tools.

The tools are built from the reference sources by ``make -f oracle/Makefile.ref`` into oracle/_ref/
(this container only; /root/reference is not needed afterwards). Every fixture written here is data:
inputs and the reference's outputs.

  c1.*            dsearch_desert 0 200000 P 8           (200,000 seed indices, CPU-default gate)
  r3m.*           dsearch_desert 0 3000000 P 8          (larger range: exact step-8 floods)
  rg1m.*          dsearch_desert 0 0 P 8 stdin=1 < gated_1e12_1m.txt (seeds gated with the GPU-gate
                  rules over k in [1e12, 1e12 + 2^20); the list itself is an input fixture)
  st1m.*          the same with study=1 (no exact floods; st1m.zones.csv.gz holds the sorted per-zone
                  feature rows of PREFIX.zones.csv; `python make_goldens.py study` rewrites only these)
  rgx.*           the 92 gated seeds with exact_min=0 refine=0 verify=0: every proxy start of the two best
                  zones gets the step-8 and the step-4 flood and a CANDIDATE line (dsearch_desert.cpp:498-531)
  fx.*            production Params on the seeds of fx_seeds.txt (the seeds of desert_finds.jsonl, then two
                  seeds whose CANDIDATE lines a GPU run printed): step-4 floods and CANDIDATE lines at the
                  production thresholds. `python make_goldens.py exact` rewrites only rgx.* and fx.*
  TAG.cand        the sorted CANDIDATE lines the run printed on stderr (one per line)
  dflood.jsonl    dflood SEED X Z STEP at the verified desert finds, steps 8 and 4 (all rows) and
                  step 1 (the first rows)
  biomemap.npz    biomemap SEED X Z HALF STEP at verified desert finds (n x n uint8 maps; `python
                  make_goldens.py biomemap` rewrites only this file)
  dfunc_S_P.txt   dfunc S P 8 rows (proxy-table histograms) for (64, 300) and the Makefile's `make dtable`
                  size (2000, 2000); `python make_goldens.py dfunc` rewrites only these
Sorted CSV rows are stored (the reference writes rows in thread completion order).
"""
import concurrent.futures as cf
import gzip
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")


def run_search(tag, args, stdin=None):
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, tag)
        err = subprocess.run([os.path.join(REF, "dsearch_desert")] + [args[0], args[1], pre] + args[2:], check=True,
                             stdin=open(stdin) if stdin else None, stderr=subprocess.PIPE, text=True).stderr
        cand = sorted(l for l in err.splitlines() if l.startswith("CANDIDATE "))
        rows = sorted(open(pre + ".csv").read().splitlines()[1:])
        srows = sorted(open(pre + ".seeds.csv").read().splitlines()[1:])
        js = json.load(open(pre + ".json"))
        zrows = sorted(open(pre + ".zones.csv").read().splitlines()[1:]) if os.path.exists(pre + ".zones.csv") else None
    if zrows is not None:
        with gzip.GzipFile(os.path.join(HERE, tag + ".zones.csv.gz"), "wb", mtime=0) as f:
            f.write(("\n".join(zrows) + "\n").encode())
    keep = {k: js[k] for k in ("start", "done", "passed", "scanned", "sites")}
    keep["seeds_rows"] = len(srows)
    keep["candidates"] = len(cand)
    keep["seeds_sha256"] = hashlib.sha256("\n".join(srows).encode()).hexdigest()
    with open(os.path.join(HERE, tag + ".csv"), "w") as f:
        f.write("\n".join(rows) + "\n")
    if len(srows) <= 1000:
        with open(os.path.join(HERE, tag + ".seeds.csv"), "w") as f:
            f.write("\n".join(srows) + "\n")
    if cand:
        with open(os.path.join(HERE, tag + ".cand"), "w") as f:
            f.write("\n".join(cand) + "\n")
    with open(os.path.join(HERE, tag + ".json"), "w") as f:
        json.dump(keep, f, indent=1)
    print(tag, keep)


def dflood(seed, x, z, step):
    out = subprocess.run([os.path.join(REF, "dflood"), str(seed), str(x), str(z), str(step)], capture_output=True,
                         text=True)
    m = re.search(r"step (\d+): (\d+) cells", out.stdout)
    if not m:
        return {"seed": seed, "x": x, "z": z, "step": step, "cells": None, "edge": None}
    return {"seed": seed, "x": x, "z": z, "step": step, "cells": int(m.group(2)) // (step * step),
            "edge": "window edge" in out.stdout}


BIOMEMAP_CASES = [(0, 2000, 16), (1, 2000, 16), (2, 6000, 40), (5, 1200, 8)]   # (find row, half, step)


def biomemap_goldens():
    import numpy as np
    finds = [json.loads(l) for l in open(os.path.join(HERE, "desert_finds.jsonl"))]
    arrs = {}
    for row, half, step in BIOMEMAP_CASES:
        r = finds[row]
        args = [r["seed"], str(r["x"]), str(r["z"]), str(half), str(step)]
        out = subprocess.run([os.path.join(REF, "biomemap")] + args, capture_output=True, check=True).stdout
        n = int(np.frombuffer(out[:4], np.int32)[0])
        arrs["map_%d_%d_%d" % (row, half, step)] = np.frombuffer(out[4:], np.uint8).reshape(n, n)
        arrs["args_%d_%d_%d" % (row, half, step)] = np.array([int(a) for a in args], dtype=object).astype(str)
    np.savez_compressed(os.path.join(HERE, "biomemap.npz"), **arrs)
    print("biomemap", sorted(arrs))


DFUNC_CASES = [(64, 300), (2000, 2000)]


def dfunc_goldens():
    for seeds, pts in DFUNC_CASES:
        out = subprocess.run([os.path.join(REF, "dfunc"), str(seeds), str(pts), "8"], capture_output=True, text=True,
                             check=True).stdout
        with open(os.path.join(HERE, "dfunc_%d_%d.txt" % (seeds, pts)), "w") as f:
            f.write(out)
        print("dfunc", seeds, pts, len(out.splitlines()))


def study_goldens():
    run_search("st1m", ["0", "0", "8", "stdin=1", "study=1"], stdin=os.path.join(HERE, "gated_1e12_1m.txt"))


def fx_seed_list():
    """The seeds of the verified finds (signed, as desert_finds.jsonl holds them), in file order, then the
    two extra seeds; duplicates dropped."""
    seeds = []
    for line in open(os.path.join(HERE, "desert_finds.jsonl")):
        r = json.loads(line)
        if "seed" in r and "x" in r:
            seeds.append(int(r["seed"]))
    seeds += [1368900756516249850, 2136669149758385146]
    out = []
    for s in seeds:
        if s not in out:
            out.append(s)
    return out


def exact_goldens():
    run_search("rgx", ["0", "0", "8", "stdin=1", "exact_min=0", "refine=0", "verify=0"],
               stdin=os.path.join(HERE, "gated_1e12_1m.txt"))
    with open(os.path.join(HERE, "fx_seeds.txt"), "w") as f:
        f.write("".join("%d\n" % s for s in fx_seed_list()))
    run_search("fx", ["0", "0", "8", "stdin=1"], stdin=os.path.join(HERE, "fx_seeds.txt"))


def main():
    if sys.argv[1:2] == ["exact"]:
        return exact_goldens()
    if sys.argv[1:2] == ["biomemap"]:
        return biomemap_goldens()
    if sys.argv[1:2] == ["dfunc"]:
        return dfunc_goldens()
    if sys.argv[1:2] == ["study"]:
        return study_goldens()
    run_search("c1", ["0", "200000", "8"])
    run_search("r3m", ["0", "3000000", "8"])
    run_search("rg1m", ["0", "0", "8", "stdin=1"], stdin=os.path.join(HERE, "gated_1e12_1m.txt"))
    finds = [json.loads(l) for l in open(os.path.join(HERE, "desert_finds.jsonl"))]
    pts = [(int(r["seed"]), r["x"], r["z"]) for r in finds if "x" in r and "seed" in r]
    jobs = [(s, x, z, st) for st in (8, 4) for (s, x, z) in pts] + [(s, x, z, 1) for (s, x, z) in pts[:12]]
    res = []
    with cf.ThreadPoolExecutor(int(sys.argv[1]) if len(sys.argv) > 1 else 6) as ex:
        for r in ex.map(lambda j: dflood(*j), jobs):
            res.append(r)
    with open(os.path.join(HERE, "dflood.jsonl"), "w") as f:
        for r in res:
            f.write(json.dumps(r) + "\n")
    print("dflood", len(res))
    biomemap_goldens()
    dfunc_goldens()
    study_goldens()
    exact_goldens()


if __name__ == "__main__":
    main()
