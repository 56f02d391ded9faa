import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "swift-high-performance-sparse-dense-matrix-multiplication-on-gpus_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: full-size configuration tests")


def _ensure_built():
    need = [os.path.join(PKG, "libswift_amd.so"), os.path.join(ROOT, "oracle", "liboracle.so")]
    if not all(os.path.exists(p) for p in need):
        import __graft_entry__
        __graft_entry__.build()


@pytest.fixture(scope="session")
def swift():
    _ensure_built()
    import swift_amd
    swift_amd.lib()
    return swift_amd


@pytest.fixture(scope="session")
def oracle():
    _ensure_built()
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def cuda(swift):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch.cuda.is_available() is False")
    torch.cuda.init()
    return torch.device("cuda:0")
