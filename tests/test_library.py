"""The C-ABI library: it loads on a machine without a GPU and exports every
function include/te.h declares (no compute calls here)."""
import ctypes
import os
import re

import te_host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "te.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(te_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(te_host.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_kernel_registry_covers_the_oracle():
    import oracle_lib
    mine = set(te_host.kernel_names())
    assert "_copy" in mine
    assert set(oracle_lib.kernel_names()) <= mine


def test_config_defaults_roundtrip():
    c = te_host.config()
    assert (c.nx, c.ny, c.nz) == (64, 64, 40)
    assert c.nsound == 6 and abs(c.dt - 6.0) < 1e-6
    c2 = te_host.config(nx=32, tke=1)
    assert c2.nx == 32 and c2.tke == 1 and c2.ny == 64


def test_program_create_needs_a_gpu_or_rejects_unknown_kernels():
    """te_program_create validates names before touching the device."""
    import numpy as np
    import pytest
    with pytest.raises(te_host.TeError, match="no kernel"):
        te_host.Replayer(1024, ["not_a_kernel"], np.zeros((1, 64), np.uint32))
