"""The FP64 CPU oracle (oracle.cpp), loaded through ctypes.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it as the checker. The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_open.restype = ctypes.c_void_p
        L.orc_open.argtypes = [ctypes.c_void_p]
        L.orc_close.argtypes = [ctypes.c_void_p]
        L.orc_sdf.argtypes = [ctypes.c_void_p, _dp, ctypes.c_size_t, _dp, _dp, _u32p, _u32p, ctypes.c_int]
        L.orc_bake.restype = ctypes.c_long
        L.orc_bake.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _u32p, _u32p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_long), ctypes.c_int]
        L.orc_brick.restype = ctypes.c_int
        L.orc_brick.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.POINTER(ctypes.c_int16)]
        L.orc_decide.restype = ctypes.c_long
        L.orc_decide.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_size_t, ctypes.c_int]
        L.orc_counts.restype = ctypes.c_int
        L.orc_counts.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.c_int]
        L.orc_face_inside.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_double]
        L.orc_face_uv.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _dp, _dp]
        L.orc_face_foot.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _dp, _dp]
        L.orc_nurbs_closest.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _dp, ctypes.c_double, _dp, _dp]
        L.orc_nurbs_eval.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_edge_closest.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _dp, _dp, _dp]
        _lib = L
    return _lib


# The primitives the oracle counts (oracle.h ORC_COUNT_*), in order.
COUNT_NAMES = ("patch_eval", "surface_newton_step", "curve_eval", "curve_newton_step", "foot_plane", "foot_cylinder", "foot_cone",
               "foot_sphere", "foot_torus", "uv_plane", "uv_cylinder", "uv_cone", "uv_sphere", "uv_torus", "trim_crossing_step",
               "edge_line", "edge_arc", "vertex")


def counts(reset=True):
    """The primitives every oracle call executed since the last reset (summed over its
    threads), as {name: count}; reset them when `reset`."""
    out = (ctypes.c_uint64 * len(COUNT_NAMES))()
    n = lib().orc_counts(out, len(COUNT_NAMES), 1 if reset else 0)
    assert n == len(COUNT_NAMES), "oracle counts out of step with COUNT_NAMES"
    return dict(zip(COUNT_NAMES, (int(v) for v in out)))


def _p(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    """FP64 exact signed distance of a part (restating solid3::Exact)."""

    def __init__(self, part):
        self.part = part
        self.h = lib().orc_open(ctypes.cast(part.desc, ctypes.c_void_p))
        if not self.h:
            raise ValueError("the oracle takes surfaces and spline edges up to degree 7")

    def sdf(self, xyz, bound=None, threads=0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = len(xyz)
        d = np.empty(n)
        feat = np.empty(n, dtype=np.uint32)
        flags = np.empty(n, dtype=np.uint32)
        b = None if bound is None else np.ascontiguousarray(bound, dtype=np.float64)
        lib().orc_sdf(self.h, _p(xyz), n, None if b is None else _p(b), _p(d), feat.ctypes.data_as(_u32p), flags.ctypes.data_as(_u32p), threads)
        return d, feat, flags

    def bake(self, level, threads=0, atlas_cap=None):
        nb = level.nb
        n = nb[0] * nb[1] * nb[2]
        cap = atlas_cap if atlas_cap is not None else n * 365
        m = np.empty(n, dtype=np.uint32)
        atlas = np.empty(max(cap, 1), dtype=np.uint32)
        read = ctypes.c_long(0)
        slots = lib().orc_bake(self.h, ctypes.byref(level), m.ctypes.data_as(_u32p), atlas.ctypes.data_as(_u32p), cap, ctypes.byref(read), threads)
        if slots < 0:
            raise RuntimeError("atlas capacity too small")
        return m, atlas[: slots * 365].copy(), read.value

    def decide(self, level, threads=0):
        nb = level.nb
        out = np.empty(nb[0] * nb[1] * nb[2], dtype=np.int64)
        k = lib().orc_decide(self.h, ctypes.byref(level), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(out), threads)
        return out[:k].copy()

    def brick(self, first_node, h):
        q = np.empty(729, dtype=np.int16)
        fn = np.ascontiguousarray(first_node, dtype=np.float64)
        cls = lib().orc_brick(self.h, _p(fn), h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)))
        return cls, q

    def face_inside(self, face, u, v):
        return bool(lib().orc_face_inside(self.h, face, u, v))

    def face_uv(self, face, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        uv = np.empty(2)
        lib().orc_face_uv(self.h, face, _p(q), _p(uv))
        return uv

    def nurbs_eval(self, face, u, v):
        out = np.empty(3)
        lib().orc_nurbs_eval(self.h, face, u, v, _p(out))
        return out

    def nurbs_closest(self, face, p, cutoff=float("inf")):
        p = np.ascontiguousarray(p, dtype=np.float64)
        d = np.empty(1)
        q = np.empty(3)
        ok = lib().orc_nurbs_closest(self.h, face, _p(p), cutoff, _p(d), _p(q))
        return (float(d[0]), q) if ok else None

    def edge_closest(self, edge, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        d = np.empty(1)
        q = np.empty(3)
        interior = lib().orc_edge_closest(self.h, edge, _p(p), _p(d), _p(q))
        return float(d[0]), q, bool(interior)

    def close(self):
        if self.h:
            lib().orc_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
