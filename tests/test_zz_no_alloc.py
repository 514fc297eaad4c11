"""te_region_hashes_noalloc: region hashes of a caller's device memory with
no allocation (accumulators in the library's device array), against
te_region_hash of a context holding the same words (MI355X)."""
import ctypes
import sys

import numpy as np
import pytest

import te_host

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(dptr, regions):
    L = te_host.lib()
    n = len(regions)
    offs = (ctypes.c_uint64 * max(n, 1))(*[o for o, _ in regions])
    cnts = (ctypes.c_uint64 * max(n, 1))(*[c for _, c in regions])
    out = np.zeros(max(n, 1), np.uint32)
    rc = L.te_region_hashes_noalloc(ctypes.c_void_p(dptr), offs, cnts, n,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return rc, out[:n]


def test_noalloc_hashes_match_te_region_hash(gpu):
    import torch
    rng = np.random.default_rng(3)
    words = rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    dev = torch.from_numpy(words.view(np.int32)).to("cuda:0")
    torch.cuda.synchronize()
    ctx = te_host.Replayer(words.size, [], np.zeros((0, 64), np.uint32))
    ctx.upload(0, words)
    calls = [[(0, 1 << 20), (5, 1000), (777, 65536), (1 << 19, 3), (1000, 0)] + [(i * 4099, 4097) for i in range(200)],
             [(17, 12345), (0, 1)]]  # the second call is smaller: fewer and shorter regions
    for q, regions in enumerate(calls):
        say("call", q, "regions", len(regions))
        rc, got = _noalloc(dev.data_ptr(), regions)
        assert rc == te_host.TE_OK
        want = np.array([ctx.region_hash(o, c) for o, c in regions], np.uint32)
        assert np.array_equal(got, want), np.nonzero(got != want)[0][:5]
    ctx.close()
    # beyond the accumulator array: rejected, nothing launched
    rc, _ = _noalloc(dev.data_ptr(), [(0, 4)] * 257)
    assert rc == te_host.TE_ERR_ARG
    rc, _ = _noalloc(dev.data_ptr(), [(2 ** 64 - 8, 16)])
    assert rc == te_host.TE_ERR_ARG
