import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


def _ensure_oracle_built():
    so = ROOT / "oracle" / "_build" / "liboracle.so"
    if not so.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True,
                       stdout=subprocess.DEVNULL)


_ensure_oracle_built()


@pytest.fixture(scope="session")
def oracle():
    from oracle_ctypes import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle_ctypes import Reference, reference_available
    if not reference_available():
        pytest.skip("reference numerics library oracle/_ref was not built "
                    "(reference tree absent)")
    return Reference()


_torch_device_ready = False


@pytest.fixture(autouse=True)
def _torch_device_first(request):
    """GPU tests move planes through torch tensors.  torch's HIP runtime is
    initialised before the library's own (which a solver created first would
    do), once per session."""
    global _torch_device_ready
    if request.node.get_closest_marker("gpu") is not None and not _torch_device_ready:
        import torch
        assert torch.cuda.is_available(), "GPU test without a visible HIP device"
        torch.cuda.init()
        _torch_device_ready = True
    yield
