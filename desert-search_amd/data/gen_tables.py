#!/usr/bin/env python3
"""Regenerate ``mc_tables.inc`` (Minecraft 26.3 overworld climate data) from the reference checkout.

The biome-parameter tree, its parameter intervals, the tree's level strides and the terrain-offset
spline are Minecraft's own data, exported from Cubiomes (MIT) into the reference's
``src/worldgen_data.hpp``; the 40x40 desert-probability table is the reference's measured proxy
table (``src/desert_dtable.hpp``).  Both must be bit-identical to what the reference evaluates, so
they are carried over as data, re-emitted here in this package's own layout:

* ``DSX_SPLINE``   spline nodes {axis, n, value, x[12], slope[12], child[12]}
* ``DSX_PARAM``    parameter intervals [lo, hi] (10000 x climate value)
* ``DSX_TREE``     packed tree nodes (6 x 8-bit parameter index | child/biome | leaf marker)
* ``DSX_STEPS``    per-level child strides of the packed tree
* ``DSX_DPROB``    P(desert | C_lo bin, E_lo bin), bins of 0.05

Usage: python gen_tables.py [/root/reference] [--dtable desert_dtable.hpp]   (writes mc_tables.inc beside
this script; --dtable takes the proxy table from a header written by gen_dtable.py)
"""
import os
import re
import struct
import sys

ARGS = sys.argv[1:]
DTABLE = None
if "--dtable" in ARGS:   # a proxy table regenerated with dfunc + gen_dtable.py instead of the reference's
    i = ARGS.index("--dtable")
    DTABLE = ARGS[i + 1]
    del ARGS[i:i + 2]
REF = ARGS[0] if ARGS else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def body(text, name):
    """Text between the opening '{' after `name` and its matching '};'."""
    i = text.index(name)
    i = text.index("{", text.index("=", i))
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[i + 1:j]
    raise ValueError(name)


def top_level_items(s):
    items, depth, cur = [], 0, ""
    for ch in s:
        if ch == "{":
            depth += 1
        if ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            if cur.strip():
                items.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        items.append(cur.strip())
    return items


def fl(tok):
    """Parse a C float literal (hex or decimal, optional f suffix) -> exact python float."""
    tok = tok.strip()
    if tok.endswith("f") and not tok.lower().startswith("0x") or (tok.lower().startswith(("0x", "-0x")) and tok.endswith("f")):
        tok = tok[:-1]
    if "x" in tok.lower():
        return float.fromhex(tok)
    return float(tok)


def main():
    wg = open(os.path.join(REF, "src", "worldgen_data.hpp")).read()
    dt = open(DTABLE or os.path.join(REF, "src", "desert_dtable.hpp")).read()

    splines = []
    for it in top_level_items(body(wg, "splines[]")):
        inner = it.strip()[1:-1]
        parts = top_level_items(inner)
        axis, n, value = int(parts[0]), int(parts[1]), fl(parts[2])

        def arr(p, conv):
            p = p.strip()[1:-1].strip()
            return [conv(v) for v in p.split(",")] if p else []

        xs, slopes, child = arr(parts[3], fl), arr(parts[4], fl), arr(parts[5], int)
        splines.append((axis, n, value, xs, slopes, child))

    params = []
    for it in top_level_items(body(wg, "parameters[][2]")):
        a, b = it.strip()[1:-1].split(",")
        params.append((int(a), int(b)))

    tree = [int(t.strip().rstrip("ULL"), 16) for t in top_level_items(body(wg, "tree[]"))]
    steps = [int(t) for t in top_level_items(body(wg, "steps[]"))]

    prob = []
    for row in top_level_items(body(dt, "desert_pass_prob[40][40]")):
        prob.append([fl(v) for v in row.strip()[1:-1].split(",")])
    assert len(prob) == 40 and all(len(r) == 40 for r in prob)

    def f32(v):
        v = struct.unpack("<f", struct.pack("<f", v))[0]       # the literal's float32 value
        return float.hex(v) + "f" if v != 0 else "0.0f"

    out = []
    out.append("// Generated by gen_tables.py from the reference's src/worldgen_data.hpp and src/desert_dtable.hpp.")
    out.append("// Minecraft 26.3 overworld biome-parameter tree + terrain spline: data exported from Cubiomes (MIT,")
    out.append("// (c) 2020 Cubitect) via the reference; desert proxy table: the reference's measured table (MIT).")
    out.append("// Data only. Included by the HIP library and by the CPU oracle.")
    out.append("#pragma once")
    out.append("#include <stdint.h>")
    out.append("typedef struct { int axis, n; float value; float x[12], slope[12]; int child[12]; } dsx_spline_node;")
    out.append(f"#define DSX_NSPLINE {len(splines)}")
    out.append(f"#define DSX_NPARAM {len(params)}")
    out.append(f"#define DSX_NTREE {len(tree)}")
    out.append(f"#define DSX_NSTEPS {len(steps)}")
    out.append("static const dsx_spline_node DSX_SPLINE[DSX_NSPLINE] = {")
    for axis, n, value, xs, slopes, child in splines:
        def pad(v, k, conv):
            v = list(v) + [0] * (k - len(v))
            return "{" + ",".join(conv(x) for x in v) + "}"
        out.append("  {%d,%d,%s,%s,%s,%s}," % (axis, n, f32(value), pad(xs, 12, f32), pad(slopes, 12, f32),
                                             pad(child, 12, str)))
    out.append("};")
    out.append("static const int32_t DSX_PARAM[DSX_NPARAM][2] = {")
    for i in range(0, len(params), 8):
        out.append("  " + ",".join("{%d,%d}" % p for p in params[i:i + 8]) + ",")
    out.append("};")
    out.append("static const uint64_t DSX_TREE[DSX_NTREE] = {")
    for i in range(0, len(tree), 6):
        out.append("  " + ",".join("0x%016xull" % t for t in tree[i:i + 6]) + ",")
    out.append("};")
    out.append("static const int DSX_STEPS[DSX_NSTEPS] = {" + ",".join(map(str, steps)) + "};")
    out.append("static const float DSX_DPROB[40][40] = {")
    for r in prob:
        out.append("  {" + ",".join(f32(v) for v in r) + "},")
    out.append("};")
    with open(os.path.join(HERE, "mc_tables.inc"), "w") as f:
        f.write("\n".join(out) + "\n")
    print(f"splines {len(splines)} params {len(params)} tree {len(tree)} steps {steps}")


if __name__ == "__main__":
    main()
