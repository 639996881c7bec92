import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session", autouse=True)
def _built_oracle():
    """The oracle is compiled from source on first use (gcc, seconds)."""
    import oracle_lib
    if not os.path.exists(os.path.join(oracle_lib.ORACLE_DIR, "liboracle.so")):
        oracle_lib.build()
    yield


def gpu_available():
    try:
        import torch  # noqa: F401  (device count only; does not initialise HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def rift():
    """The product library; a GPU test fails loudly when it is not built."""
    from rift_amd import kp
    kp.load()
    return kp


@pytest.fixture
def progress(capsys):
    """Progress lines of a long test, written past pytest's output capture
    so that a run that watches its output sees the test working."""
    def say(msg):
        with capsys.disabled():
            print(msg, flush=True)
    return say
