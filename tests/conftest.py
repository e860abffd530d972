import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_lib():
    """The C-ABI library, loaded on a machine with a GPU.  Fails loudly if it is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from elastic_fm_vp_vs_rho_amd import _capi
    return _capi.lib()
