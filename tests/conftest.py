import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def built():
    """The native libraries, built in-tree when missing (host compilers only)."""
    from implicitcam_amd import _native

    need = [os.path.join(_native.LIBDIR, "libicam_gen.so"), os.path.join(ROOT, "oracle", "_build", "liboracle.so")]
    if not all(os.path.exists(p) for p in need):
        import subprocess

        from implicitcam_amd import build

        build.build_gen()
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    return True


@pytest.fixture(scope="module", autouse=True)
def _torch_gpu_first(request):
    """Modules of GPU tests (some hand torch's device memory to the library) start
    torch's HIP runtime before any fixture opens a handle: torch finds no GPU when the
    library's own HIP calls came first in the process."""
    if request.node.get_closest_marker("gpu") is not None:
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
    yield
