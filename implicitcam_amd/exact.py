"""Host mirror of the reference's GPU runner `PartGpu` (crates/icam-sim/src/partgpu.rs)
and of the device half of its brick cache `PartCache` (crates/icam-sim/src/partcache.rs),
over the C ABI of libicam_hip.so (include/icam_hip.h).

Method names and meanings follow `PartGpu`: `points` / `points_refined` take
(point, clamp) records and give (value, flag) pairs, `bricks` takes (first node, h)
records and gives 729 (value, flag) pairs each; flags 0 exact, 1 a bound past the
clamp, 2 undecided (partgpu.rs:227-264, gpu.wgsl:1913-1922), 3 left to the caller (a
host feature, one the device does not read, came within the answer: the reference's
Code::Cpu, crates/icam-core/src/solid3/gpu.rs:714-716). Errors raise
`RuntimeError` with the library's message, as `PartGpu` returns Err(String).
`query` and `bake` are the additive entry points: unbounded signed distances, and
a whole brick level (map and atlas in the layout of crates/icam-sweep/src/cache.rs:74-101;
bricks left to the caller are HOST entries, read on the CPU as partcache.rs:794-797 does).
"""
import ctypes

import numpy as np

from . import _native

FP64 = 0
MIXED = 1

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class BakeStats(ctypes.Structure):
    _fields_ = [("bricks", ctypes.c_int64), ("band", ctypes.c_int64), ("slots", ctypes.c_int64), ("undecided", ctypes.c_int64),
                ("refined", ctypes.c_int64), ("unsure_trim", ctypes.c_int64), ("unsure_tie", ctypes.c_int64), ("settled", ctypes.c_int64),
                ("overflowed", ctypes.c_int64), ("host", ctypes.c_int64)]


AIR, SOLID, HOST = 0xFFFFFFFF, 0xFFFFFFFE, 0xFFFFFFFD  # map entries that are not slots


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _native._load("libicam_hip.so")
        L.icam_hip_open.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]
        L.icam_hip_close.argtypes = [ctypes.c_void_p]
        L.icam_hip_last_error.restype = ctypes.c_char_p
        L.icam_hip_device_name.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.icam_hip_warm.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.icam_hip_nodes.restype = ctypes.c_size_t
        L.icam_hip_nodes.argtypes = [ctypes.c_void_p]
        L.icam_hip_points.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, _dp]
        L.icam_hip_bricks.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, _dp]
        L.icam_hip_query.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t, _dp, _u32p, _u32p]
        L.icam_hip_query_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]
        L.icam_hip_query_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_void_p]
        L.icam_hip_bake.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, _u32p, _u32p, ctypes.c_size_t, ctypes.POINTER(BakeStats)]
        L.icam_hip_bake_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                           ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(BakeStats), ctypes.c_void_p]
        L.icam_hip_stage_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), _dp, ctypes.c_int]
        L.icam_hip_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.icam_hip_cast.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.icam_hip_query_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(BakeStats)]
        L.icam_hip_batch_open.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]
        L.icam_hip_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.icam_hip_batch_sync.argtypes = [ctypes.c_void_p]
        L.icam_hip_batch_result.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(BakeStats)]
        L.icam_hip_batch_download.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u32p, _u32p, ctypes.c_size_t, ctypes.POINTER(BakeStats)]
        L.icam_hip_batch_tuning.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.icam_hip_batch_stage_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), _dp, ctypes.c_int]
        L.icam_hip_batch_close.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().icam_hip_last_error().decode())


def cast_device(levels, d_maps, d_atlases, d_rays, n, d_out, d_hit, stream=None):
    """The viewer's ray cast over baked levels (icam_hip_cast): `levels` icam_level records
    finest first, their device maps and atlases (ints, as bake_device or a batch's result
    gives them), n rays at d_rays (6 doubles each: origin, unit direction); d_out gets 4
    floats a ray (t, -1 for none; the gradient there), d_hit 2 words (level, slot)."""
    k = len(levels)
    lv = (_native.Level * k)(*levels)
    maps = (ctypes.c_void_p * k)(*d_maps)
    atl = (ctypes.c_void_p * k)(*d_atlases)
    _check(lib().icam_hip_cast(lv, maps, atl, k, d_rays, n, d_out, d_hit, stream))


class PartGpu:
    """The part's exact signed distance on the GPU (PartGpu::new, partgpu.rs:132-186).

    `need`: the distance (mm) within which the B-spline sub-patch and curve-piece
    trees certify a single minimum (the reference packs for 4 finest cells times
    sqrt 3, crates/icam-sim/src/main.rs:644)."""

    def __init__(self, part, need):
        self.part = part
        self.need = need
        h = ctypes.c_void_p()
        _check(lib().icam_hip_open(ctypes.cast(part.desc, ctypes.c_void_p), need, ctypes.byref(h)))
        self.h = h
        buf = ctypes.create_string_buffer(256)
        _check(lib().icam_hip_device_name(self.h, buf, 256))
        self.adapter = buf.value.decode()
        self.nodes = lib().icam_hip_nodes(self.h)

    def warm(self, secs):
        _check(lib().icam_hip_warm(self.h, secs))

    def points(self, pts, clamp, precision=FP64):
        """(value, flag) per (point, clamp) record."""
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
        rec = np.empty((len(pts), 4))
        rec[:, :3] = pts
        rec[:, 3] = np.broadcast_to(np.asarray(clamp, dtype=np.float64), (len(pts),))
        rec = np.ascontiguousarray(rec)
        out = np.empty((len(pts), 2))
        _check(lib().icam_hip_points(self.h, precision, rec.ctypes.data, len(pts), out.ctypes.data_as(_dp)))
        return out

    def points_refined(self, pts, clamp):
        return self.points(pts, clamp, MIXED)

    def bricks(self, firsts, h, precision=FP64):
        """729 (value, flag) pairs per brick whose first node is firsts[i], cell side h."""
        firsts = np.asarray(firsts, dtype=np.float64).reshape(-1, 3)
        rec = np.empty((len(firsts), 4))
        rec[:, :3] = firsts
        rec[:, 3] = np.broadcast_to(np.asarray(h, dtype=np.float64), (len(firsts),))
        rec = np.ascontiguousarray(rec)
        out = np.empty((len(firsts), 729, 2))
        _check(lib().icam_hip_bricks(self.h, precision, rec.ctypes.data, len(firsts), out.ctypes.data_as(_dp)))
        return out

    def query(self, xyz, precision=FP64):
        """Signed distances, flags and nearest features at world points (unbounded)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = len(xyz)
        d = np.empty(n)
        fl = np.empty(n, dtype=np.uint32)
        ft = np.empty(n, dtype=np.uint32)
        _check(lib().icam_hip_query(self.h, precision, xyz.ctypes.data_as(_dp), n, d.ctypes.data_as(_dp), fl.ctypes.data_as(_u32p),
                                    ft.ctypes.data_as(_u32p)))
        return d, fl, ft

    def query_device(self, d_xyz, n, d_d, d_flags, d_feat=None, precision=FP64, stream=None):
        """The same on device pointers (ints), enqueued on `stream` (a hipStream_t as int)."""
        _check(lib().icam_hip_query_device(self.h, precision, d_xyz, n, d_d, d_flags, d_feat, stream))

    def query_device_noalloc(self, d_xyz, n, d_d, d_flags, d_feat=None, precision=FP64, stream=None):
        """query_device with no allocation (scratch from the library's static array, at
        most 2^18 points a call)."""
        _check(lib().icam_hip_query_device_noalloc(self.h, precision, d_xyz, n, d_d, d_flags, d_feat, stream))

    def query_stats(self):
        """Statistics of the last query: points searched again in double precision (for
        a trim decision near a curve, a near-tie) and points settled (waits)."""
        st = BakeStats()
        _check(lib().icam_hip_query_stats(self.h, ctypes.byref(st)))
        return st

    def bake(self, level, precision=FP64):
        """The level's brick map and atlas (PartCache::build on the device), and statistics."""
        nb = level.nb
        n = nb[0] * nb[1] * nb[2]
        m = np.empty(n, dtype=np.uint32)
        st = BakeStats()
        cap = n * 365
        atlas = np.empty(cap, dtype=np.uint32)
        _check(lib().icam_hip_bake(self.h, precision, ctypes.byref(level), m.ctypes.data_as(_u32p), atlas.ctypes.data_as(_u32p), cap,
                                   ctypes.byref(st)))
        return m, atlas[: st.slots * 365].copy(), st

    def bake_device(self, level, precision=FP64, stream=None, stats=False):
        """The level baked into device memory; returns (map pointer, atlas pointer, stats or None)."""
        dm, da = ctypes.c_void_p(), ctypes.c_void_p()
        st = BakeStats() if stats else None
        _check(lib().icam_hip_bake_device(self.h, precision, ctypes.byref(level), ctypes.byref(dm), ctypes.byref(da),
                                          ctypes.byref(st) if st is not None else None, stream))
        return dm.value, da.value, st

    def tune(self, level, precision=FP64):
        """The single-level bake's autotuner (icam_hip_tune): lanes a node of the nodes'
        search timed on this level, the fastest kept for later bakes in `precision`;
        returns its report."""
        buf = ctypes.create_string_buffer(512)
        _check(lib().icam_hip_tune(self.h, precision, ctypes.byref(level), buf, 512))
        return buf.value.decode()

    def stage_times(self):
        """Device time (ms) of each kernel stage of the last bake or query."""
        names = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_double * 32)()
        k = lib().icam_hip_stage_times(self.h, names, ms, 32)
        return {names[i].decode(): ms[i] for i in range(k)}

    def close(self):
        if getattr(self, "h", None):
            lib().icam_hip_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many parts' brick levels baked by one launch of a captured hipGraph (the
    additive batch entry points of include/icam_hip.h): `parts` are PartGpu handles,
    `levels` their icam_level records. `tune`: run the autotuner over the persistent
    grid sizes of the queue stages first."""

    def __init__(self, parts, levels, precision=MIXED, tune=True):
        n = len(parts)
        self.parts = list(parts)
        self.levels = (_native.Level * n)(*levels)
        hs = (ctypes.c_void_p * n)(*[p.h.value for p in parts])
        b = ctypes.c_void_p()
        _check(lib().icam_hip_batch_open(hs, ctypes.cast(self.levels, ctypes.c_void_p), n, precision, 1 if tune else 0, ctypes.byref(b)))
        self.b = b
        self.n = n

    def run(self, stream=None):
        _check(lib().icam_hip_batch_run(self.b, stream))

    def sync(self):
        _check(lib().icam_hip_batch_sync(self.b))

    def stats(self, i):
        st = BakeStats()
        _check(lib().icam_hip_batch_result(self.b, i, None, None, ctypes.byref(st)))
        return st

    def download(self, i):
        """Part i's map, atlas and statistics (host arrays)."""
        lv = self.levels[i]
        nb = lv.nb[0] * lv.nb[1] * lv.nb[2]
        st = self.stats(i)
        m = np.empty(nb, dtype=np.uint32)
        atlas = np.empty(max(st.slots * 365, 1), dtype=np.uint32)
        _check(lib().icam_hip_batch_download(self.b, i, m.ctypes.data_as(_u32p), atlas.ctypes.data_as(_u32p), len(atlas), ctypes.byref(st)))
        return m, atlas[: st.slots * 365].copy(), st

    def stage_times(self):
        """Device time (ms) of each stage of one run outside the graph (waits)."""
        names = (ctypes.c_char_p * 48)()
        ms = (ctypes.c_double * 48)()
        k = lib().icam_hip_batch_stage_times(self.b, names, ms, 48)
        if k < 0:
            _check(1)
        return {names[i].decode(): ms[i] for i in range(k)}

    @property
    def tuning(self):
        buf = ctypes.create_string_buffer(4096)
        _check(lib().icam_hip_batch_tuning(self.b, buf, 4096))
        return buf.value.decode()

    def close(self):
        if getattr(self, "b", None):
            lib().icam_hip_batch_close(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
