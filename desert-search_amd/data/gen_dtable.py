#!/usr/bin/env python3
"""Proxy table from the rows of the ``dfunc`` tool (the reference's tools/gen_dtable.py:1-17 step).

    desert-search_amd/dfunc SEEDS POINTS THREADS > dfunc.txt
    python gen_dtable.py dfunc.txt > desert_dtable.hpp
    python gen_tables.py --dtable desert_dtable.hpp      # adopt it into mc_tables.inc

Rows are "set ci ei total desert passable". Set 1 (C octaves 0-2, E octaves 0-1) is used:
P[ci][ei] = passable / total, printed with 3 decimals, 0 for cells with fewer than 20 samples. The
output is a C++ header in the reference's src/desert_dtable.hpp layout (namespace standalone,
``desert_pass_prob[40][40]``), so either build can include it.
"""
import sys

NB = 40
MIN_SAMPLES = 20


def table(lines):
    """(hot points of set 1, 40 x 40 list of (total, passable))."""
    cells = [[[0, 0] for _ in range(NB)] for _ in range(NB)]
    for line in lines:
        f = line.split()
        if len(f) != 6 or f[0] != "1":
            continue
        ci, ei, total, passable = int(f[1]), int(f[2]), int(f[3]), int(f[5])
        cells[ci][ei][0] += total
        cells[ci][ei][1] += passable
    return sum(c[0] for row in cells for c in row), cells


def header(lines):
    npts, cells = table(lines)
    out = ["// P(desert or bridgeable river | T>=0.55, C octaves 0-2, E octaves 0-1), measured on "
           "%.1fM hot points" % (npts / 1e6),
           "// (tools/dfunc.cpp, tools/gen_dtable.py). Bins of 0.05: C_lo in [-0.8,1.2) (rows), E_lo in [-1,1) "
           "(cols); <20 samples -> 0.",
           "#pragma once", "namespace standalone {", "inline constexpr float desert_pass_prob[40][40] = {"]
    for row in cells:
        vals = ["%.3ff" % (p / t if t >= MIN_SAMPLES else 0) for t, p in row]
        out.append("{" + ",".join(vals) + "},")
    out += ["};", "}"]
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    if len(sys.argv) != 2:
        sys.exit("usage: gen_dtable.py dfunc.txt > desert_dtable.hpp")
    sys.stdout.write(header(open(sys.argv[1]).read().splitlines()))
