"""ctypes bindings of the native libraries built in-tree.

`libicam_hip.so` is the product: the C ABI declared in include/icam_hip.h over
the HIP kernels. `libicam_gen.so` makes the synthetic parts (the evaluator's
input data model, include/icam_part.h). Loading fails loudly when a library has
not been built: there is no fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "_lib")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)


class PartDesc(ctypes.Structure):
    """The head of `icam_part_desc` (include/icam_part.h): counts, arrays, box, tol."""

    _fields_ = [
        ("n_faces", ctypes.c_uint32),
        ("n_edges", ctypes.c_uint32),
        ("n_verts", ctypes.c_uint32),
        ("faces", ctypes.c_void_p),
        ("edges", ctypes.c_void_p),
        ("verts", ctypes.c_void_p),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("tol", ctypes.c_double),
    ]


class Level(ctypes.Structure):
    """`icam_level`: bricks of 8^3 cells of side h from the node lo, nb per axis."""

    _fields_ = [("lo", ctypes.c_double * 3), ("h", ctypes.c_double), ("nb", ctypes.c_int32 * 3)]


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() (python -c 'import __graft_entry__ as g; g.build()')")
    return ctypes.CDLL(path)


_gen = None


def gen():
    global _gen
    if _gen is None:
        lib = _load("libicam_gen.so")
        lib.icam_gen_part.restype = ctypes.POINTER(PartDesc)
        lib.icam_gen_part.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
        lib.icam_gen_free.argtypes = [ctypes.POINTER(PartDesc)]
        lib.icam_gen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_size_t, c_double_p, c_double_p, c_double_p]
        lib.icam_gen_near_trim.argtypes = [ctypes.POINTER(PartDesc), ctypes.c_uint64, ctypes.c_size_t, ctypes.c_double, c_double_p]
        lib.icam_gen_near_surface.argtypes = [ctypes.POINTER(PartDesc), ctypes.c_uint64, ctypes.c_size_t, ctypes.c_double, c_double_p]
        lib.icam_gen_describe.restype = ctypes.c_size_t
        lib.icam_gen_describe.argtypes = [ctypes.POINTER(PartDesc), c_double_p, ctypes.c_size_t]
        _gen = lib
    return _gen
