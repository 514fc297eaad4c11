"""Test configuration: the `gpu` marker, import paths and built libraries.

CPU tests (`-m "not gpu"`) cover the oracle against the reference's golden
vectors, the encoder and host logic, and that the C-ABI library loads and
exports every declared symbol.  GPU tests (`-m gpu`) call the HIP path
through the C ABI and compare it with the oracle and the golden vectors.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "thunderengine-web_amd")
for p in (ROOT, os.path.join(ROOT, "tests"), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _ensure_built():
    """Builds the oracle and libte.so when they are absent (never rebuilds)."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "_build", "liboracle.so")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if not os.path.exists(os.path.join(PKG, "lib", "libte.so")):
        subprocess.check_call(["make", "-s", "-j8", "-C", PKG])


_ensure_built()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("GPU test run without a visible GPU")
    import torch
    torch.cuda.init()
    return torch.device("cuda:0")
