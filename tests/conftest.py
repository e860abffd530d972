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


@pytest.fixture(autouse=True, scope="module")
def _free_kept_states():
    """After each test module, free the forward states that the differentiable
    op keeps between calls (a full-size C2 tape holds ~160 GB), so that a later
    module's strategy budget sees the whole GPU."""
    yield
    ops = sys.modules.get("elastic_fm_vp_vs_rho_amd.ops")
    if ops is not None:
        ops.release_state_pools()
        import gc
        import torch
        gc.collect()
        if torch.cuda.is_available():
            torch.cuda.empty_cache()
