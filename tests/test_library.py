"""The C-ABI library: it loads on a machine without a GPU and exports every
function include/te.h declares (no compute calls here)."""
import ctypes
import os
import re

import numpy as np

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


def test_workgroup_renumbering_division_is_exact():
    """te_block() (csrc/te_device.hpp) divides workgroup numbers l < 2^24 by
    grid extents d <= 2^16 as (l * ceil(2^40 / d)) >> 40.  Checked by brute
    force against integer division on every d up to 1024 and on the largest
    extents, at l = 0, the ends of the range and every multiple of d +- 1
    near them."""
    def magic(d):
        return -(-(1 << 40) // d)

    ds = list(range(1, 1025)) + [4095, 4096, 4097, 40000, 65535, 65536]
    for d in ds:
        m = magic(d)
        assert m < 1 << 41
        ls = {0, 1, (1 << 24) - 1, (1 << 24) - 2}
        for base in (d, 2 * d, ((1 << 24) - 1) // d * d, ((1 << 24) - 1) // d * d - d):
            ls.update({base - 1, base, base + 1})
        for l in ls:
            if 0 <= l < 1 << 24:
                assert (l * m) >> 40 == l // d, (d, l)
    # exhaustive over l for a few extents the bench grids use
    for d in (3, 4, 48, 50, 192):
        m = magic(d)
        l = np.arange(1 << 24, dtype=np.uint64)
        assert np.array_equal((l * np.uint64(m)) >> np.uint64(40), l // np.uint64(d))


def test_autotune_rejects_missing_contexts():
    """te_autotune and its queries on no context: an error code (no GPU call)."""
    L = te_host.lib()
    assert L.te_autotune(None, 5, 0) != te_host.TE_OK
    assert L.te_tuned_ops(None) == 0
    a, b = ctypes.c_float(), ctypes.c_float()
    assert L.te_autotune_ms(None, ctypes.byref(a), ctypes.byref(b)) != te_host.TE_OK


def test_word_ranges_that_wrap_are_rejected():
    """Offsets near 2^64 (a negative offset from a caller) would wrap
    `off + n`: the C ABI rejects them with TE_ERR_ARG before any device
    call (an encoded-only context, no GPU needed)."""
    run = te_host.Run(te_host.config(nx=8, ny=8, nz=4), encode_only=True)
    L = te_host.lib()
    buf = np.zeros(16, np.uint32)
    ptr = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    for off in (2**64 - 1, 2**64 - 8, run.mem_words):
        assert L.te_upload_words(run.h, ctypes.c_uint64(off), ptr, ctypes.c_uint64(16)) == te_host.TE_ERR_ARG
        assert L.te_download_words(run.h, ctypes.c_uint64(off), ptr, ctypes.c_uint64(16)) == te_host.TE_ERR_ARG
    assert L.te_download_words(run.h, ctypes.c_uint64(0), ptr, ctypes.c_uint64(2**64 - 1)) == te_host.TE_ERR_ARG
    h = ctypes.c_uint32()
    assert L.te_region_hash(run.h, ctypes.c_uint64(2**64 - 1), ctypes.c_uint64(2), ctypes.byref(h)) == te_host.TE_ERR_ARG
    assert L.te_fnv1a(run.h, ctypes.c_uint64(2**64 - 1), ctypes.c_uint64(2), ctypes.byref(h)) == te_host.TE_ERR_ARG
    # an in-range read of the encoded image still works
    assert L.te_download_words(run.h, ctypes.c_uint64(0), ptr, ctypes.c_uint64(16)) == te_host.TE_OK
    run.close()
