"""The C ABI: the library loads and exports every entry point its headers declare
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared("icam_hip.h", "icam_hip_")
    for n in ["icam_hip_open", "icam_hip_close", "icam_hip_last_error", "icam_hip_device_name", "icam_hip_warm", "icam_hip_points",
              "icam_hip_bricks", "icam_hip_query", "icam_hip_query_device", "icam_hip_bake", "icam_hip_bake_device", "icam_hip_stage_times",
              "icam_hip_nodes", "icam_hip_tune"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from implicitcam_amd import _native

    path = os.path.join(_native.LIBDIR, "libicam_hip.so")
    if not os.path.exists(path):
        pytest.skip("libicam_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for n in declared("icam_hip.h", "icam_hip_"):
        assert hasattr(lib, n), n


def test_generator_exports(built):
    from implicitcam_amd import _native

    lib = _native.gen()
    for n in ["icam_gen_part", "icam_gen_free", "icam_gen_uniform"]:
        assert hasattr(lib, n)


def test_last_error_without_a_device():
    """With no device, opening fails loudly with a message (PartGpu::new returns Err)."""
    from implicitcam_amd import _native

    path = os.path.join(_native.LIBDIR, "libicam_hip.so")
    if not os.path.exists(path):
        pytest.skip("libicam_hip.so not built")
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from implicitcam_amd import Part, PartGpu

    with pytest.raises(RuntimeError, match="device"):
        PartGpu(Part("can"), 1.0)
