import ctypes
import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_package():
    """Import desert-search_amd/ (not a Python identifier) as `desert_search_amd`."""
    if "desert_search_amd" in sys.modules:
        return sys.modules["desert_search_amd"]
    path = os.path.join(REPO, "desert-search_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("desert_search_amd", path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["desert_search_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class FloodOut(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_uint64), ("sampled", ctypes.c_uint64), ("minx", ctypes.c_int32),
                ("maxx", ctypes.c_int32), ("minz", ctypes.c_int32), ("maxz", ctypes.c_int32),
                ("edge", ctypes.c_int32), ("border", ctypes.c_int32)]

    def tuple(self):
        return (self.cells, self.sampled, self.minx, self.maxx, self.minz, self.maxz, bool(self.edge), bool(self.border))


class Oracle:
    """ctypes view of oracle/liboracle.so (the CPU checker)."""

    def __init__(self):
        path = os.path.join(REPO, "oracle", "liboracle.so")
        if not os.path.exists(path):
            pytest.skip("oracle/liboracle.so not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        u64, f32, f64, i32, i64, vp = (ctypes.c_uint64, ctypes.c_float, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int64, ctypes.c_void_p)
        sig = {
            "or_mix": (u64, [u64]),
            "or_phase": (f64, [u64, vp]),
            "or_gate_kern": (None, [i32, f32, vp]),
            "or_gate_amp": (f32, [u64, i32, i32, i32, f32]),
            "or_gate_side": (f32, [u64, i32, i32, i32, i32, f32]),
            "or_gate_grid": (None, [u64, i32, i32, i32, vp]),
            "or_gate_blur_max": (f32, [vp, i32, i32, vp]),
            "or_tabulate_side": (None, [u64, i32, f64, vp]),
            "or_perlin_side": (f64, [u64, i32, f64, f64]),
            "or_gate_run": (i64, [u64, u64, f32, f32, f32, i32, vp, vp, i64, vp]),
            "or_phase_list": (i64, [u64, u64, f32, vp, i64]),
            "or_amplitude_sub": (f32, [u64]),
            "or_tables": (None, [u64, f64, f64, i32, vp]),
            "or_scan": (i64, [u64, f64, f64, f64, f64, vp, i64, vp]),
            "or_noise": (f64, [u64, i32, f64, f64, f64]),
            "or_noise_batch": (None, [u64, i32, vp, vp, i64, vp]),
            "or_climate": (None, [u64, i32, i32, vp]),
            "or_pred_batch": (None, [u64, i32, f32, vp, vp, i64, vp]),
            "or_flood": (None, [u64, i32, f32, i32, i32, i32, i32, i32, vp, vp]),
            "or_run": (i32, [u64, u64, vp, i64, i32, i32, ctypes.c_char_p, vp, vp, vp, i32]),
            "or_biome": (i32, [u64, i32, i32]),
            "or_biomemap": (i32, [u64, i32, i32, i32, i32, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    def __getattr__(self, name):
        return getattr(self.L, name)

    # convenience wrappers
    def gate_run(self, start, count, maxph=0.06, pthr=0.645, amp=0.70, threads=8, cap=1 << 16):
        idx = np.zeros(cap, np.uint64)
        amps = np.zeros(cap, np.float32)
        cnt = np.zeros(3, np.uint64)
        n = self.L.or_gate_run(start, count, maxph, pthr, amp, threads, idx.ctypes.data, amps.ctypes.data, cap,
                               cnt.ctypes.data)
        return idx[:n], amps[:n], tuple(int(c) for c in cnt)

    def gate_kern(self, R, sig):
        k = np.zeros(2 * R + 1, np.float32)
        self.L.or_gate_kern(R, sig, k.ctypes.data)
        return k

    def gate_grid(self, seed, side, G, box):
        V = np.zeros((G, G), np.float32)
        self.L.or_gate_grid(seed, side, G, box, V.ctypes.data)
        return V

    def biomemap(self, seed, X, Z, half, step):
        """(n, n x n uint8 map) as tools/biomemap.cpp writes it; n < 0 when the centre is not desert."""
        m = 2 * half // step + 1
        out = np.zeros((m, m), np.uint8)
        n = self.L.or_biomemap(seed, X, Z, half, step, out.ctypes.data)
        return n, out

    def flood(self, seed, kind, cut, qx, qz, step, half, conn8):
        out = FloodOut()
        self.L.or_flood(seed, kind, cut, qx, qz, step, half, conn8, ctypes.byref(out), None)
        return out.tuple()

    def pred_batch(self, seed, kind, xs, zs, cut=0.35):
        x = np.ascontiguousarray(xs, dtype=np.int32)
        z = np.ascontiguousarray(zs, dtype=np.int32)
        out = np.zeros(len(x), np.int8)
        self.L.or_pred_batch(seed, kind, cut, x.ctypes.data, z.ctypes.data, len(x), out.ctypes.data)
        return out

    def phase_list(self, start, count, maxph, cap=1 << 22):
        out = np.zeros(cap, np.uint64)
        n = self.L.or_phase_list(start, count, maxph, out.ctypes.data, cap)
        return out[:n]


@pytest.fixture(autouse=True, scope="module")
def flood_health(request):
    """After every GPU test module: no flood batch reached the engine's time limit (and was run again) or
    exhausted its tile pool, counted process-wide by dsx_flood_stats."""
    yield
    marks = getattr(request.module, "pytestmark", [])
    marks = marks if isinstance(marks, list) else [marks]
    pkg = sys.modules.get("desert_search_amd")
    if not any(m.name == "gpu" for m in marks) or pkg is None or pkg.dsx._lib is None:
        return
    st = pkg.dsx.FloodStats()
    assert pkg.dsx._lib.dsx_flood_stats_get(ctypes.byref(st), 0) == 0
    assert (st.time_retries, st.capacity_resets) == (0, 0), \
        f"flood engine: {st.time_retries} time-limit retries, {st.capacity_resets} capacity resets"


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def engine(pkg):
    eng = pkg.dsx.Engine(0)
    yield eng
    eng.close()
