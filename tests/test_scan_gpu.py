"""Stages S1-S2 on the GPU (period tables, blur, maxta/maxtb, raw scan candidates) against the CPU
oracle, bit for bit and in the CPU's scan order."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_package

pytestmark = pytest.mark.gpu

GATED = [int(l.split()[0]) % (1 << 64) for l in open(os.path.join(GOLDEN, "gated_1e12_1m.txt"))]
RAW = np.dtype([("s", "<f4"), ("t", "<f4"), ("c", "<f4"), ("e", "<f4"), ("x", "<i4"), ("z", "<i4")])


def oracle_scan(oracle, seed, ta=0.25, tt=0.55, wc=0.25, we=0.41):
    buf = np.zeros(1 << 17, RAW)
    mt = np.zeros(2, np.float32)
    n = oracle.or_scan(seed, ta, tt, wc, we, buf.ctypes.data, len(buf), mt.ctypes.data)
    return mt, buf[:n]


@pytest.mark.parametrize("sigma,sigmac", [(2.0, 2.0), (1.2, 2.4), (1.0, 1.5)])
def test_tables_bitexact(engine, oracle, sigma, sigmac):
    dsx = load_package().dsx
    for seed in GATED[:2]:
        got = engine.tables(seed, dsx.ScanParams(sigma=sigma, sigmac=sigmac))
        want = np.zeros((6, 512, 512), np.float32)
        oracle.or_tables(seed, sigma, sigmac, 1, want.ctypes.data)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_scan_bitexact_gated_seeds(engine, oracle):
    """The 92 GPU-gated seeds of the golden list: maxta/maxtb and every raw candidate, in order."""
    seeds = np.array(GATED[:48], np.uint64)
    maxta, maxtb, cands = engine.scan(seeds)
    for k, s in enumerate(seeds):
        mt, want = oracle_scan(oracle, int(s))
        assert maxta[k].view(np.uint32) == mt[0].view(np.uint32)
        assert maxtb[k].view(np.uint32) == mt[1].view(np.uint32)
        got = cands[k]
        assert len(got) == len(want)
        assert np.array_equal(got.view(np.uint32).reshape(-1, 6), want.view(np.uint32).reshape(-1, 6))


def test_scan_chunking_and_capacity(engine, oracle):
    """More seeds than one table chunk (64); a too-small cap reports the true counts."""
    dsx = load_package().dsx
    seeds = np.array((GATED * 2)[:80], np.uint64)
    maxta, maxtb, cands = engine.scan(seeds)
    for k in (0, 63, 64, 79):
        assert np.array_equal(cands[k].view(np.uint32), cands[k % len(GATED)].view(np.uint32))
    n = np.array([len(c) for c in cands])
    with pytest.raises(dsx.DsxError) as e:
        engine.scan(seeds[:4], cap=8)
    assert e.value.code == dsx.DSX_E_CAPACITY


def test_scan_other_thresholds(engine, oracle):
    dsx = load_package().dsx
    p = dsx.ScanParams(ta=0.2, tt=0.5, wc=0.3, we=0.35)
    seeds = np.array(GATED[:4], np.uint64)
    maxta, maxtb, cands = engine.scan(seeds, p, cap=1 << 18)
    for k, s in enumerate(seeds):
        mt, want = oracle_scan(oracle, int(s), 0.2, 0.5, 0.3, 0.35)
        assert np.array_equal(cands[k].view(np.uint32).reshape(-1, 6), want.view(np.uint32).reshape(-1, 6))
