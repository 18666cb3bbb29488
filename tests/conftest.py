import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an MI355X (gfx950) and libvoxgpu.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def voxlib():
    """libvoxgpu.so, built in-tree when absent (CPU compile)."""
    from voxvision_amd import _lib, build
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    return _lib.load()


@pytest.fixture(autouse=True)
def _announce(request):
    """Name every GPU case on stderr before it touches the device."""
    if request.node.get_closest_marker("gpu") is not None:
        sys.stderr.write("\n[gpu-case] %s\n" % request.node.nodeid)
        sys.stderr.flush()
    yield


@pytest.fixture(scope="session")
def gpu():
    """Device 0 checked to be gfx950, and the library loaded."""
    import torch
    from voxvision_amd import _lib
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    L = _lib.load()
    _lib.check(L.voxg_init(0), "voxg_init")
    torch.zeros(1, device="cuda")
    return L
