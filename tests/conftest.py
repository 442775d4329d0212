import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_built():
    """Tests load what is built and build only when nothing is."""
    lib = os.path.join(ROOT, "umpalumpa_amd", "libumpalumpa_amd.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    cpp = [os.path.join(ROOT, "tests", "cpp", "_build", name)
           for name in ("fr_fp_tests", "fr_harness")]
    if not (os.path.exists(lib) and os.path.exists(orc) and all(map(os.path.exists, cpp))):
        import __graft_entry__
        __graft_entry__.build()


_ensure_built()


@pytest.fixture(autouse=True)
def _hip_runtime_first(request):
    """GPU tests bring the HIP runtime up before their body runs: in a process that first
    allocated tens of GB of host memory (the oracle's box-1024 volumes), the runtime's first
    initialisation found no device (hipErrorNoDevice) on this image."""
    if request.node.get_closest_marker("gpu") is not None:
        import torch
        torch.cuda.init()
        torch.zeros(1, device="cuda")
    yield


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def reference():
    import pyoracle
    if not pyoracle.Reference.available():
        pytest.skip("reference CPU code not compiled here (oracle/_ref)")
    return pyoracle.Reference()


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load
