import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "slow: full-size case (seconds of GPU time)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (test infrastructure), built on first use."""
    from oracle import oracle as O

    O.build()
    O.lib()
    return O


_gpu_checked = None


def _gpu_available():
    global _gpu_checked
    if _gpu_checked is None:
        try:
            import torch

            _gpu_checked = bool(torch.cuda.is_available())
        except Exception:
            _gpu_checked = False
    return _gpu_checked


@pytest.fixture(scope="session")
def tofe():
    """The HIP library binding; GPU tests need a device."""
    if not _gpu_available():
        pytest.skip("no GPU device visible")
    from topopt3d_amd import _lib

    _lib.lib()
    return _lib
