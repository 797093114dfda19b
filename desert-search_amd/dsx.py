"""Python binding of libdsx.so (include/dsx.h) - the host-side mirror of the reference's GPU engine.

The reference exposes its GPU path as C++ functions in ``src/gpu/gproxy.h`` and ``src/gpu/gate.cuh``
(``tg::run_gate``, ``gproxy::start_gate``, the ``tgate`` tool's ``calib`` mode). This module binds the
C ABI that replaces them, with the same argument meaning, via ctypes. It only loads the in-tree
``libdsx.so``: if the library is missing or no gfx950 device is present, it raises - there is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsx.so")

DSX_OK, DSX_E_ARG, DSX_E_HIP, DSX_E_CAPACITY, DSX_E_NODEVICE, DSX_E_BUSY, DSX_E_INTERRUPTED = 0, 1, 2, 3, 4, 5, 6
DSX_NOALLOC_MAX_COUNT, DSX_NOALLOC_MAX_PHASE = 1 << 22, 1 << 16   # dsx_gate_noalloc limits per call
HW_QUEUES = 20   # hardware queues requested from the HIP runtime (pipeline workers, the gate; more measured slower)
HW_QUEUES_MAX = 28   # the most measured since the flood engine's stream-ordering fix (host/hwq.h)


def hw_queues():
    """GPU_MAX_HW_QUEUES for this process: DSX_HW_QUEUES if set, clamped to [1, HW_QUEUES_MAX], else 20."""
    v = os.environ.get("DSX_HW_QUEUES", "")
    q = int(v) if v.strip() else HW_QUEUES
    return max(1, min(HW_QUEUES_MAX, q))


class DsxError(RuntimeError):
    def __init__(self, code, what, detail=""):
        self.code = code
        super().__init__(f"{what}: {detail or 'error'} (code {code})")


class GateParams(ctypes.Structure):
    """dsx_gate_params; defaults are the reference's production GPU gate (gproxy.cu:409, gate.cuh:349-355
    with dsearch_desert.cpp:76 gate_phase/gate_amp/gate_pthr)."""
    _fields_ = [("max_phase", ctypes.c_float), ("amp", ctypes.c_float), ("pthr", ctypes.c_float),
                ("margin", ctypes.c_float), ("sig", ctypes.c_float), ("box", ctypes.c_int)]

    def __init__(self, max_phase=0.06, amp=0.70, pthr=0.645, margin=0.01, sig=1.0, box=1):
        super().__init__(max_phase, amp, pthr, margin, sig, box)


class ScanParams(ctypes.Structure):
    """dsx_scan_params; defaults are the reference's Params (dsearch_desert.cpp:53-58)."""
    _fields_ = [("ta", ctypes.c_float), ("tt", ctypes.c_float), ("wc", ctypes.c_float), ("we", ctypes.c_float),
                ("sigma", ctypes.c_double), ("sigmac", ctypes.c_double)]

    def __init__(self, ta=0.25, tt=0.55, wc=0.25, we=0.41, sigma=2.0, sigmac=2.0):
        super().__init__(ta, tt, wc, we, sigma, sigmac)


RAW_CAND = np.dtype([("s", "<f4"), ("t", "<f4"), ("c", "<f4"), ("e", "<f4"), ("x", "<i4"), ("z", "<i4")])


FLOOD_JOB = np.dtype([("seed", "<u8"), ("kind", "<i4"), ("cut", "<f4"), ("qx", "<i4"), ("qz", "<i4"), ("step", "<i4"),
                      ("half", "<i4"), ("conn8", "<i4"), ("stop_cells", "<i4")])
AREA = np.dtype([("cells", "<u8"), ("sampled", "<u8"), ("minx", "<i4"), ("maxx", "<i4"), ("minz", "<i4"),
                 ("maxz", "<i4"), ("edge", "<i4"), ("border", "<i4")])
PRED_DESERT_STATE, PRED_HOT, PRED_PROXY, PRED_DESERT = 0, 1, 2, 3
POINT_T, POINT_PROXYPROB, POINT_HOTTEST, POINT_BIOME, POINT_PROXY_C, POINT_PROXY_E = 16, 17, 18, 19, 20, 21
TILE_MASK = np.dtype([("job", "<u4"), ("tx", "<i4"), ("tz", "<i4"), ("rows", "<u4", (32,))])
AREA_QUERY = np.dtype([("seed", "<u8"), ("x", "<i4"), ("z", "<i4"), ("half", "<i4"), ("pad", "<i4")])
FLOOD_HINT = np.dtype([("job", "<u4"), ("tx", "<i4"), ("tz", "<i4"), ("blocks", "<u4")])


class SearchParams(ctypes.Structure):
    """dsx_search_params = the reference's struct Params (dsearch_desert.cpp:50-78), same defaults."""
    _fields_ = [(n, ctypes.c_double) for n in ("max_phase", "amp", "amp_margin", "sigma", "sigmac", "wc", "we", "t", "ta")] + \
        [("max_sites", ctypes.c_int32), ("pcut", ctypes.c_double), ("hot_min", ctypes.c_double), ("multi", ctypes.c_int32),
         ("exact_n", ctypes.c_int32), ("exact_min", ctypes.c_double), ("hot_pre", ctypes.c_double),
         ("csv_min", ctypes.c_double), ("nms", ctypes.c_int32), ("refine_mb", ctypes.c_double),
         ("verify_mb", ctypes.c_double)] + \
        [(n, ctypes.c_int32) for n in ("study", "gpu", "gpucheck", "gscan", "gscancheck", "gate")] + \
        [(n, ctypes.c_double) for n in ("gate_phase", "gate_amp", "gate_pthr")] + [("from_stdin", ctypes.c_int32),
                                                                               ("workers", ctypes.c_int32),
                                                                               ("prepass", ctypes.c_int32),
                                                                               ("mcverify", ctypes.c_int32)]


class SearchCounts(ctypes.Structure):
    _fields_ = [("done", ctypes.c_uint64), ("passed", ctypes.c_uint64), ("scanned", ctypes.c_uint64),
                ("sites", ctypes.c_uint64), ("hot_zones", ctypes.c_uint64), ("big_zones", ctypes.c_uint64),
                ("sec_tables", ctypes.c_double), ("sec_scan", ctypes.c_double), ("sec_sites", ctypes.c_double),
                ("sec_proxy", ctypes.c_double), ("sec_hot", ctypes.c_double), ("sec_gpu", ctypes.c_double)]


SEARCH_ENUM, SEARCH_GATE, SEARCH_LIST = 0, 1, 2


class ServerConfig(ctypes.Structure):
    _fields_ = [("workers", ctypes.c_int32), ("area_batch", ctypes.c_int32), ("hbm_budget", ctypes.c_uint64),
                ("params", SearchParams)]


class Area(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("cells", "sampled")] + \
        [(n, ctypes.c_int32) for n in ("minx", "maxx", "minz", "maxz", "edge", "border")]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("kind", ctypes.c_int32), ("status", ctypes.c_int32),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("area", Area), ("counts", SearchCounts), ("csv_bytes", ctypes.c_uint64),
                ("seeds_bytes", ctypes.c_uint64)]


def flood_jobs(rows):
    """rows of (seed, kind, cut, qx, qz, step, half, conn8) -> FLOOD_JOB array"""
    a = np.zeros(len(rows), FLOOD_JOB)
    for k, r in enumerate(rows):
        a[k] = tuple(r) + (0,)
    return a


class FloodStats(ctypes.Structure):
    """dsx_flood_stats: process-wide flood-engine counters (per kind index DSX_PRED_*)."""
    _fields_ = [("calls", ctypes.c_uint64), ("rounds", ctypes.c_uint64), ("floods", ctypes.c_uint64 * 4),
                ("tiles", ctypes.c_uint64 * 4), ("evaluated", ctypes.c_uint64 * 4), ("sampled_cold", ctypes.c_uint64 * 4),
                ("sampled_hot", ctypes.c_uint64 * 4), ("small_floods", ctypes.c_uint64),
                ("small_sampled", ctypes.c_uint64), ("eval_ms", ctypes.c_double), ("eval_launches", ctypes.c_uint64),
                ("size_floods", (ctypes.c_uint64 * 6) * 4), ("size_tiles", (ctypes.c_uint64 * 6) * 4),
                ("size_sampled", (ctypes.c_uint64 * 6) * 4), ("eval_rounds", ctypes.c_uint64),
                ("eval_needed", ctypes.c_uint64), ("wg_wait", ctypes.c_uint64), ("wg_eval", ctypes.c_uint64),
                ("wg_load", ctypes.c_uint64), ("wg_life", ctypes.c_uint64), ("time_retries", ctypes.c_uint64),
                ("capacity_resets", ctypes.c_uint64)]


class GateCounts(ctypes.Structure):
    _fields_ = [("phase", ctypes.c_uint64), ("pre", ctypes.c_uint64), ("full", ctypes.c_uint64)]


_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)

_lib = None


def load(path=LIB_PATH):
    """Load libdsx.so (once). Raises if the built library is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DsxError(DSX_E_NODEVICE, "libdsx.so not built", path)
    # hardware queues for the pipeline worker streams, the gate and spares, read by the HIP runtime when it starts
    os.environ["GPU_MAX_HW_QUEUES"] = str(hw_queues())
    lib = ctypes.CDLL(path)
    lib.dsx_strerror.restype = ctypes.c_char_p
    lib.dsx_strerror.argtypes = [ctypes.c_int]
    lib.dsx_last_error.restype = ctypes.c_char_p
    lib.dsx_last_error.argtypes = []
    lib.dsx_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.dsx_destroy.argtypes = [ctypes.c_void_p]
    lib.dsx_gate.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(GateParams), _u64p, _f32p,
                             _u64p, ctypes.c_uint64, _u64p, ctypes.POINTER(GateCounts)]
    lib.dsx_gate_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.POINTER(GateParams), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_uint64, ctypes.c_void_p]
    lib.dsx_set_gate_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.dsx_gate_amps.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64, ctypes.POINTER(GateParams), _f32p, _f32p]
    lib.dsx_gate_profile.argtypes = [ctypes.c_void_p, _f64p, _u64p]
    lib.dsx_scan.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64, ctypes.POINTER(ScanParams), _f32p, _f32p,
                             ctypes.c_void_p, ctypes.c_uint64, _u64p]
    lib.dsx_tables.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ScanParams), _f32p]
    lib.dsx_noise.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                              ctypes.c_void_p]
    lib.dsx_climate.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                ctypes.c_void_p]
    lib.dsx_predicates.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    lib.dsx_flood.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    lib.dsx_flood_hinted.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                                     ctypes.c_void_p]
    lib.dsx_flood_masks.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, _u64p]
    lib.dsx_points.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                               ctypes.c_void_p]
    lib.dsx_search_defaults.argtypes = [ctypes.POINTER(SearchParams)]
    lib.dsx_search_defaults.restype = None
    lib.dsx_search.argtypes = [ctypes.c_void_p, ctypes.POINTER(SearchParams), ctypes.c_int, ctypes.c_uint64,
                               ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_char_p,
                               ctypes.POINTER(SearchCounts)]
    lib.dsx_area_queries.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    lib.dsx_area_workload.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    lib.dsx_area_workload.restype = None
    lib.dsx_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
    lib.dsx_server_destroy.argtypes = [ctypes.c_void_p]
    lib.dsx_server_workers.argtypes = [ctypes.c_void_p]
    lib.dsx_server_pool_clear.argtypes = [ctypes.c_int]
    lib.dsx_submit_scan.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, _u64p]
    lib.dsx_submit_area.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u64p]
    lib.dsx_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Result)]
    lib.dsx_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Result)]
    lib.dsx_result_text.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p,
                                    ctypes.c_uint64]
    lib.dsx_release.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.dsx_profile.argtypes = [ctypes.c_int]
    lib.dsx_set_precision.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.dsx_noise_f32.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                  ctypes.c_void_p]
    lib.dsx_flood_stats_get.argtypes = [ctypes.POINTER(FloodStats), ctypes.c_int]
    lib.dsx_dfunc.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    for f in ("dsx_create", "dsx_destroy", "dsx_gate", "dsx_gate_amps", "dsx_gate_profile", "dsx_scan", "dsx_tables",
              "dsx_noise", "dsx_climate", "dsx_predicates", "dsx_flood", "dsx_flood_masks", "dsx_points", "dsx_search", "dsx_area_queries",
              "dsx_server_create", "dsx_server_destroy", "dsx_server_workers", "dsx_server_pool_clear", "dsx_submit_scan", "dsx_submit_area",
              "dsx_wait", "dsx_poll", "dsx_result_text", "dsx_release", "dsx_profile", "dsx_flood_stats_get", "dsx_set_precision", "dsx_noise_f32", "dsx_dfunc"):
        getattr(lib, f).restype = ctypes.c_int
    _lib = lib
    return lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


class Engine:
    """One device context (dsx_ctx). Not shared between threads."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        self._check(self.lib.dsx_create(device, ctypes.byref(h)), "dsx_create")
        self.h = h

    def _check(self, rc, what):
        if rc != DSX_OK:
            raise DsxError(rc, what, (self.lib.dsx_strerror(rc) + b" " + self.lib.dsx_last_error()).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.dsx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gate(self, start, count, params=None, cap=1 << 20, want_amps=True):
        """Seeds passing the GPU gate over indices [start, start+count): (seeds u64, amps f32, k u64, counts).
        want_amps=False passes no amplitude buffer (amps is then None), as the pipeline does: the gate may
        then decide the full gate from its fast amplitude."""
        p = params or GateParams()
        seeds = np.zeros(cap, np.uint64)
        amps = np.zeros(cap, np.float32) if want_amps else None
        kidx = np.zeros(cap, np.uint64)
        n = ctypes.c_uint64()
        cnt = GateCounts()
        rc = self.lib.dsx_gate(self.h, start, count, ctypes.byref(p), _ptr(seeds, ctypes.c_uint64),
                               _ptr(amps, ctypes.c_float) if want_amps else None, _ptr(kidx, ctypes.c_uint64), cap,
                               ctypes.byref(n), ctypes.byref(cnt))
        self._check(rc, "dsx_gate")
        m = n.value
        return seeds[:m], (amps[:m] if want_amps else None), kidx[:m], (cnt.phase, cnt.pre, cnt.full)

    def set_gate_graph(self, on):
        """dsx_set_gate_graph: 1 replays the gate as a hipGraph per chunk, 0 launches it stage by stage."""
        self._check(self.lib.dsx_set_gate_graph(self.h, 1 if on else 0), "dsx_set_gate_graph")

    def gate_amps(self, seeds, params=None):
        """(pre-gate amplitude, full amplitude) per seed, no thresholds (tgate calib mode)."""
        p = params or GateParams()
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        pre = np.zeros(len(s), np.float32)
        full = np.zeros(len(s), np.float32)
        self._check(self.lib.dsx_gate_amps(self.h, _ptr(s, ctypes.c_uint64), len(s), ctypes.byref(p),
                                           _ptr(pre, ctypes.c_float), _ptr(full, ctypes.c_float)), "dsx_gate_amps")
        return pre, full

    def gate_profile(self):
        ms = np.zeros(4, np.float64)
        work = np.zeros(4, np.uint64)
        self._check(self.lib.dsx_gate_profile(self.h, _ptr(ms, ctypes.c_double), _ptr(work, ctypes.c_uint64)),
                    "dsx_gate_profile")
        return ms, work

    def scan(self, seeds, params=None, cap=1 << 16):
        """Per seed: maxta, maxtb and the raw candidates (RAW_CAND records) in CPU scan order."""
        p = params or ScanParams()
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        n = len(s)
        maxta = np.zeros(n, np.float32)
        maxtb = np.zeros(n, np.float32)
        cands = np.zeros(max(n, 1) * cap, RAW_CAND)
        ncand = np.zeros(n, np.uint64)
        rc = self.lib.dsx_scan(self.h, _ptr(s, ctypes.c_uint64), n, ctypes.byref(p), _ptr(maxta, ctypes.c_float),
                               _ptr(maxtb, ctypes.c_float), cands.ctypes.data, cap, _ptr(ncand, ctypes.c_uint64))
        self._check(rc, "dsx_scan")
        return maxta, maxtb, [cands[i * cap:i * cap + int(ncand[i])] for i in range(n)]

    def tables(self, seed, params=None):
        p = params or ScanParams()
        out = np.zeros((6, 512, 512), np.float32)
        self._check(self.lib.dsx_tables(self.h, seed, ctypes.byref(p), _ptr(out, ctypes.c_float)), "dsx_tables")
        return out

    def noise(self, seed, channel, xyz):
        p = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(p), np.float64)
        self._check(self.lib.dsx_noise(self.h, seed, channel, p.ctypes.data, len(p), out.ctypes.data), "dsx_noise")
        return out

    def climate(self, seed, xs, zs):
        x = np.ascontiguousarray(xs, dtype=np.int32)
        z = np.ascontiguousarray(zs, dtype=np.int32)
        q = np.zeros((len(x), 6), np.int64)
        self._check(self.lib.dsx_climate(self.h, seed, x.ctypes.data, z.ctypes.data, len(x), q.ctypes.data),
                    "dsx_climate")
        return q

    def predicates(self, seed, kind, xs, zs, cut=0.35):
        x = np.ascontiguousarray(xs, dtype=np.int32)
        z = np.ascontiguousarray(zs, dtype=np.int32)
        out = np.zeros(len(x), np.int8)
        self._check(self.lib.dsx_predicates(self.h, seed, kind, cut, x.ctypes.data, z.ctypes.data, len(x),
                                            out.ctypes.data), "dsx_predicates")
        return out

    def dfunc(self, nseeds, pts):
        """dsx_dfunc: proxy-table histograms, uint64 array [set][ci][ei][total, desert, passable]."""
        h = np.zeros((2, 40, 40, 3), np.uint64)
        self._check(self.lib.dsx_dfunc(self.h, nseeds, pts, h.ctypes.data), "dsx_dfunc")
        return h

    def area_queries(self, queries):
        """queries: AREA_QUERY array -> AREA array (dsx_area_queries)"""
        q = np.ascontiguousarray(queries, dtype=AREA_QUERY)
        out = np.zeros(len(q), AREA)
        self._check(self.lib.dsx_area_queries(self.h, q.ctypes.data, len(q), out.ctypes.data), "dsx_area_queries")
        return out

    def flood(self, jobs):
        """jobs: FLOOD_JOB array (see flood_jobs) -> AREA array"""
        j = np.ascontiguousarray(jobs, dtype=FLOOD_JOB)
        out = np.zeros(len(j), AREA)
        self._check(self.lib.dsx_flood(self.h, j.ctypes.data, len(j), out.ctypes.data), "dsx_flood")
        return out

    def flood_hinted(self, jobs, hints):
        """dsx_flood_hinted: jobs (FLOOD_JOB array) with evaluation hints (FLOOD_HINT array) -> AREA array"""
        j = np.ascontiguousarray(jobs, dtype=FLOOD_JOB)
        h = np.ascontiguousarray(hints, dtype=FLOOD_HINT)
        out = np.zeros(len(j), AREA)
        self._check(self.lib.dsx_flood_hinted(self.h, j.ctypes.data, len(j), h.ctypes.data if len(h) else None, len(h),
                                              out.ctypes.data), "dsx_flood_hinted")
        return out

    def set_precision(self, mode):
        """dsx_set_precision: 0 plain FP64, 1 FP32 pre-pass + FP64 confirm (identical results)."""
        self._check(self.lib.dsx_set_precision(self.h, mode), "dsx_set_precision")

    def noise_f32(self, seed, channel, xyz):
        """dsx_noise_f32: the pre-pass's FP32 octaves at points xyz (n x 3) -> float64 array."""
        p = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(p), np.float64)
        self._check(self.lib.dsx_noise_f32(self.h, seed, channel, p.ctypes.data, len(p), out.ctypes.data), "dsx_noise_f32")
        return out

    def points(self, seeds, sidx, xs, zs, what, cut=0.35):
        """dsx_points: per point k of seeds[sidx[k]] at quart (xs[k], zs[k]) -> (int8 array, float64 array)."""
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        si = np.ascontiguousarray(sidx, dtype=np.uint32)
        x = np.ascontiguousarray(xs, dtype=np.int32)
        z = np.ascontiguousarray(zs, dtype=np.int32)
        o8 = np.zeros(len(x), np.int8)
        od = np.zeros(len(x), np.float64)
        self._check(self.lib.dsx_points(self.h, sd.ctypes.data, len(sd), si.ctypes.data, x.ctypes.data, z.ctypes.data,
                                        len(x), what, cut, o8.ctypes.data, od.ctypes.data), "dsx_points")
        return o8, od

    def flood_masks(self):
        """Counted cells of the last flood() call: TILE_MASK array (job, tx, tz, rows[32])."""
        n = ctypes.c_uint64()
        self.lib.dsx_flood_masks(self.h, None, 0, ctypes.byref(n))
        out = np.zeros(n.value, TILE_MASK)
        if n.value:
            self._check(self.lib.dsx_flood_masks(self.h, out.ctypes.data, n.value, ctypes.byref(n)), "dsx_flood_masks")
        return out

    def search_params(self, **kw):
        p = SearchParams()
        self.lib.dsx_search_defaults(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def search(self, prefix, mode, start=0, count=0, seeds=None, params=None):
        """Run the whole pipeline; writes PREFIX.csv / PREFIX.seeds.csv (and PREFIX.zones.csv for study=1),
        with headers, and returns the counts."""
        p = params or self.search_params()
        with open(prefix + ".csv", "w") as f:
            f.write("seed,fyA,fyB,maxta,maxtb,rank,x,z,score,t0,c0,e0,proxy,s8,s4,edge,hot\n")
        with open(prefix + ".seeds.csv", "w") as f:
            f.write("seed,fyA,fyB,maxta,maxtb,candidates,sites,best\n")
        if p.study:
            with open(prefix + ".zones.csv", "w") as f:
                f.write("seed,phT,amp,phE0,phE1,phC0,phC1,rank,x,z,t0,c0,e0,hot,proxy,f_ocean,f_badl,f_mang,f_good,"
                        "mean_c,mean_e,sd_e\n")
        s = np.ascontiguousarray(seeds if seeds is not None else [], dtype=np.uint64)
        c = SearchCounts()
        self._check(self.lib.dsx_search(self.h, ctypes.byref(p), mode, start, count, s.ctypes.data if len(s) else None,
                                        len(s), prefix.encode(), ctypes.byref(c)), "dsx_search")
        return (c.done, c.passed, c.scanned, c.sites)


def area_workload(n=4096, gen_seed=0xD5E27C4):
    """dsx_area_workload: (AREA_QUERY array, caps) of the synthetic area-query workload."""
    lib = load()
    q = np.zeros(n, AREA_QUERY)
    caps = np.zeros(n, np.uint64)
    lib.dsx_area_workload(gen_seed, n, q.ctypes.data, caps.ctypes.data)
    return q, caps


class Server:
    """dsx_server: scans and area queries submitted at any time, served by worker contexts with
    continuous batching. Thread-safe."""

    def __init__(self, device=0, workers=0, area_batch=0, hbm_budget=0, params=None):
        self.lib = load()
        cfg = ServerConfig()
        cfg.workers, cfg.area_batch, cfg.hbm_budget = workers, area_batch, hbm_budget
        if params is None:
            params = SearchParams()
            self.lib.dsx_search_defaults(ctypes.byref(params))
        cfg.params = params
        h = ctypes.c_void_p()
        self._check(self.lib.dsx_server_create(device, ctypes.byref(cfg), ctypes.byref(h)), "dsx_server_create")
        self.h = h
        self.workers = self.lib.dsx_server_workers(h)

    def _check(self, rc, what):
        if rc != DSX_OK:
            raise DsxError(rc, what, (self.lib.dsx_strerror(rc) + b" " + self.lib.dsx_last_error()).decode())

    def submit_scan(self, start, count):
        i = ctypes.c_uint64()
        self._check(self.lib.dsx_submit_scan(self.h, start, count, ctypes.byref(i)), "dsx_submit_scan")
        return i.value

    def submit_area(self, seed, x, z, half):
        q = np.zeros(1, AREA_QUERY)
        q[0] = (seed, x, z, half, 0)
        i = ctypes.c_uint64()
        self._check(self.lib.dsx_submit_area(self.h, q.ctypes.data, ctypes.byref(i)), "dsx_submit_area")
        return i.value

    def wait(self, rid):
        r = Result()
        self._check(self.lib.dsx_wait(self.h, rid, ctypes.byref(r)), "dsx_wait")
        return r

    def poll(self, rid):
        r = Result()
        rc = self.lib.dsx_poll(self.h, rid, ctypes.byref(r))
        if rc == DSX_E_BUSY:
            return None
        self._check(rc, "dsx_poll")
        return r

    def text(self, rid):
        """(csv rows, seeds rows) of a finished scan"""
        r = self.wait(rid)
        csv = ctypes.create_string_buffer(max(1, r.csv_bytes))
        seeds = ctypes.create_string_buffer(max(1, r.seeds_bytes))
        self._check(self.lib.dsx_result_text(self.h, rid, csv, r.csv_bytes, seeds, r.seeds_bytes), "dsx_result_text")
        return csv.raw[:r.csv_bytes].decode(), seeds.raw[:r.seeds_bytes].decode()

    def release(self, rid):
        self._check(self.lib.dsx_release(self.h, rid), "dsx_release")

    def close(self):
        if getattr(self, "h", None):
            self.lib.dsx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
