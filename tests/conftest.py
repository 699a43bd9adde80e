import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (str(ROOT / "mpm-lbm_amd"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libfsi_hip.so")


@pytest.fixture(scope="session")
def emu():
    """The serial CPU emulation of libfsi_hip.so (same kernel bodies, host memory)."""
    from mpm_lbm import _native

    path = ROOT / "mpm-lbm_amd" / "mpm_lbm" / "_lib" / "libfsi_emu.so"
    if not path.is_file():
        pytest.fail(f"{path} not built: run __graft_entry__.build()")
    return _native.Native(path, "emulation")


@pytest.fixture(scope="session")
def hip():
    import torch
    from mpm_lbm import _native

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return _native.load_hip()
