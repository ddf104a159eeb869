import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


def _ensure_built():
    from halotools_amd import _engine
    if not os.path.exists(_engine.LIB_PATH):
        from halotools_amd import build_ext
        build_ext.build()
    oracle_lib = os.path.join(ROOT, "oracle", "_build", "libhp_oracle.so")
    if not os.path.exists(oracle_lib):
        from oracle import mesh
        mesh.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)
