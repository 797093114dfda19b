"""Surface biome queries and the biomemap command line (tools/biomemap.cpp) on the GPU, against the
reference tool's own maps and the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO, load_package

pytestmark = pytest.mark.gpu


def test_biome_points_match_oracle(engine, oracle):
    """DSX_POINT_BIOME at scattered quarts of three seeds (every climate region, grid lookups and
    nearest-box fallbacks) equals SurfaceSampler::biome."""
    dsx = load_package().dsx
    rng = np.random.default_rng(11)
    seeds = [3008578943537042204, (1 << 64) - 6294502432331704518, 12345]
    n = 3000
    sidx = rng.integers(0, len(seeds), n).astype(np.uint32)
    xs = rng.integers(-1_875_000, 1_875_000, n).astype(np.int32)
    zs = rng.integers(-1_875_000, 1_875_000, n).astype(np.int32)
    got, _ = engine.points(seeds, sidx, xs, zs, dsx.POINT_BIOME)
    want = [oracle.or_biome(seeds[s], int(x), int(z)) for s, x, z in zip(sidx, xs, zs)]
    assert [int(b) for b in got.view(np.uint8)] == want
    assert len(set(want)) > 20


def test_biomemap_tool_matches_reference():
    """biomemap SEED X Z HALF STEP: the same n and n x n bytes as the reference tool."""
    exe = os.path.join(REPO, "desert-search_amd", "biomemap")
    d = np.load(os.path.join(GOLDEN, "biomemap.npz"))
    for key in [k for k in d.files if k.startswith("map_")]:
        args = [str(a) for a in d["args_" + key[4:]]]
        r = subprocess.run([exe] + args, capture_output=True, timeout=120)
        assert r.returncode == 0, r.stderr
        n = int(np.frombuffer(r.stdout[:4], np.int32)[0])
        assert n == d[key].shape[0]
        assert np.array_equal(np.frombuffer(r.stdout[4:], np.uint8).reshape(n, n), d[key]), key
    r = subprocess.run([exe, "1", "0", "0", "400", "16"], capture_output=True, timeout=60)
    assert r.returncode == 1 and b"centre is not desert" in r.stderr


def test_biomemap_default_size(oracle):
    """The tool's default 2001 x 2001 map (HALF 16000, STEP 16): the centre's component against the
    oracle restatement on a 4x coarser grid of the same points."""
    exe = os.path.join(REPO, "desert-search_amd", "biomemap")
    seed, x, z = 3008578943537042204, 21325314, -1975294
    r = subprocess.run([exe, str(seed), str(x), str(z)], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = int(np.frombuffer(r.stdout[:4], np.int32)[0])
    assert n == 2001
    m = np.frombuffer(r.stdout[4:], np.uint8).reshape(n, n)
    # every 4th row/column is the map of STEP 64 over the same origin, except for the component marks
    # (4-connectivity on the coarser grid differs); compare the biome ids off the component
    _, c = oracle.biomemap(seed, x, z, 16000, 64)
    sub = m[::4, ::4]
    keep = (sub != 255) & (c != 255)
    assert np.array_equal(sub[keep], c[keep])
    assert (sub == 255).sum() > 1000


@pytest.mark.parametrize("seeds,pts", [(64, 300), (2000, 2000)])
def test_dfunc_tool_matches_reference(seeds, pts):
    """dfunc SEEDS POINTS THREADS: the reference tool's rows, byte for byte (integer histograms);
    2000 x 2000 is the `make dtable` size (Makefile:58-59)."""
    exe = os.path.join(REPO, "desert-search_amd", "dfunc")
    r = subprocess.run([exe, str(seeds), str(pts), "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == open(os.path.join(GOLDEN, "dfunc_%d_%d.txt" % (seeds, pts))).read()


def test_dfunc_api(engine):
    """dsx_dfunc through the C ABI: the same histogram as the tool's rows; every hot point is counted once
    per set, desert <= passable <= total; no seeds gives an empty histogram."""
    h = engine.dfunc(64, 300)
    rows = ["%d %d %d %d %d %d" % (s, i, j, *h[s, i, j]) for s in range(2) for i in range(40) for j in range(40)
            if h[s, i, j, 0]]
    assert rows == open(os.path.join(GOLDEN, "dfunc_64_300.txt")).read().splitlines()
    assert h[0, :, :, 0].sum() == h[1, :, :, 0].sum()
    assert np.all(h[..., 1] <= h[..., 2]) and np.all(h[..., 2] <= h[..., 0])
    assert engine.dfunc(0, 300).sum() == 0
