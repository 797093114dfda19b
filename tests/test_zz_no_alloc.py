"""dsx_gate_noalloc, the gate entry point that allocates nothing, against dsx_gate through the C ABI:
survivor seeds, amplitudes (bitwise), indices and counts on consecutive calls that share the library's
device scratch array, the second call on a smaller range and another stream.

The test (the caller) owns the output buffers: it allocates them with the HIP runtime libdsx.so itself
links (torch bundles a different HIP runtime, which cannot share the process with it)."""
import ctypes

import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu

SENT = -0x5A5A5A5A5A5A5A5B   # int64 sentinel written before each call
H2D, D2H = 1, 2              # hipMemcpyKind


class Hip:
    """The few HIP runtime calls the caller side needs, from the runtime already loaded by libdsx.so."""

    def __init__(self, dsx):
        dsx.load()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        for name, args in {"hipMalloc": [ctypes.POINTER(vp), sz], "hipFree": [vp],
                           "hipMemcpy": [vp, vp, sz, ctypes.c_int], "hipStreamCreate": [ctypes.POINTER(vp)],
                           "hipStreamSynchronize": [vp], "hipStreamDestroy": [vp], "hipSetDevice": [ctypes.c_int],
                           "hipDeviceSynchronize": []}.items():
            getattr(self.L, name).argtypes = args
            getattr(self.L, name).restype = ctypes.c_int

    def ck(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def upload(self, arr):
        p = ctypes.c_void_p()
        self.ck(self.L.hipMalloc(ctypes.byref(p), arr.nbytes), "hipMalloc")
        self.ck(self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.ck(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy")
        return out

    def stream(self):
        s = ctypes.c_void_p()
        self.ck(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s


def _noalloc(dsx, hip, start, count, cap, stream, params=None):
    """Runs dsx_gate_noalloc into fresh device buffers (pre-filled with sentinels) on `stream`; returns host
    copies of the outputs."""
    lib = dsx.load()
    p = params or dsx.GateParams()
    host = [np.full(cap + 8, SENT, np.int64), np.full(cap + 8, np.nan, np.float32), np.full(cap + 8, SENT, np.int64),
            np.full(4, SENT, np.int64)]
    dev = [hip.upload(h) for h in host]
    try:
        rc = lib.dsx_gate_noalloc(0, stream, start, count, ctypes.byref(p), dev[0], dev[1], dev[2], cap, dev[3])
        assert rc == dsx.DSX_OK, (rc, lib.dsx_last_error())
        hip.ck(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
        out = [hip.download(d, h) for d, h in zip(dev, host)]
    finally:
        for d in dev:
            hip.L.hipFree(d)
    return out[0].view(np.uint64), out[1], out[2].view(np.uint64), out[3].view(np.uint64)


def _compare(engine, dsx, hip, start, count, cap, stream):
    want_s, want_a, want_k, want_c = engine.gate(start, count)
    s, a, k, c = _noalloc(dsx, hip, start, count, cap, stream)
    n = len(want_s)
    m = min(n, cap)
    print(f"dsx_gate_noalloc start={start} count={count} cap={cap}: counts {list(c)}, dsx_gate {want_c} n={n}",
          flush=True)
    assert list(c[:3]) == list(want_c) and c[3] == m
    assert np.array_equal(s[:m], want_s[:m])
    assert np.array_equal(k[:m], want_k[:m])
    assert np.array_equal(a[:m].view(np.uint32), want_a[:m].view(np.uint32))
    # nothing is written past the survivors
    assert np.all(s[m:].view(np.int64) == SENT) and np.all(k[m:].view(np.int64) == SENT)
    assert np.all(np.isnan(a[m:]))
    return n


def test_noalloc_equals_gate_on_consecutive_calls(engine):
    dsx = load_package().dsx
    hip = Hip(dsx)
    s2 = hip.stream()
    try:
        # the null stream first, then a stream of the caller's
        n1 = _compare(engine, dsx, hip, 10 ** 12, dsx.DSX_NOALLOC_MAX_COUNT, 4096, None)
        assert n1 > 100
        n2 = _compare(engine, dsx, hip, 5 * 10 ** 11 + 7, (1 << 20) + 12345, 4096, s2)
        assert 0 < n2 < n1
        # capacity below the survivor count: the first cap survivors in index order, the true count reported
        _compare(engine, dsx, hip, 10 ** 12, 1 << 21, 5, s2)
    finally:
        hip.L.hipStreamDestroy(s2)


def test_noalloc_limits():
    dsx = load_package().dsx
    hip = Hip(dsx)
    lib = dsx.load()
    p = dsx.GateParams()
    buf = hip.upload(np.zeros(16, np.int64))
    try:
        rc = lib.dsx_gate_noalloc(0, None, 0, dsx.DSX_NOALLOC_MAX_COUNT + 1, ctypes.byref(p), None, None, None, 0, buf)
        assert rc == dsx.DSX_E_ARG and b"DSX_NOALLOC_MAX_COUNT" in lib.dsx_last_error()
    finally:
        hip.L.hipFree(buf)
    assert lib.dsx_gate_noalloc(0, None, 0, 16, ctypes.byref(p), None, None, None, 0, None) == dsx.DSX_E_ARG
    # a phase cut that passes more seeds than the scratch list holds: flagged by counts[0], nothing written
    seeds, amps, kidx, counts = _noalloc(dsx, hip, 0, 1 << 18, 64, None, params=dsx.GateParams(max_phase=0.5))
    assert counts[0] > dsx.DSX_NOALLOC_MAX_PHASE and list(counts[1:]) == [0, 0, 0]
    assert np.all(seeds.view(np.int64) == SENT)
    # empty range
    seeds, amps, kidx, counts = _noalloc(dsx, hip, 123, 0, 64, None)
    assert list(counts) == [0, 0, 0, 0]
