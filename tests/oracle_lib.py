"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this.  Build the library with `make -C oracle`.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        u32p, f64p = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)
        L.oracle_run.argtypes = [u32p, ctypes.POINTER(ctypes.c_char_p), u32p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.oracle_run_f64.argtypes = [u32p, f64p, ctypes.POINTER(ctypes.c_char_p), u32p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.oracle_region_hash.argtypes = [u32p, ctypes.c_uint64]
        L.oracle_region_hash.restype = ctypes.c_uint32
        L.oracle_fnv1a.argtypes = [u32p, ctypes.c_uint64]
        L.oracle_fnv1a.restype = ctypes.c_uint32
        for fn, n in (("oracle_rr_pow", 2), ("oracle_rr_log", 1), ("oracle_rr_exp", 1), ("oracle_weno5", 6),
                      ("oracle_flx5", 5), ("oracle_flx3", 3), ("oracle_upstrpd", 4)):
            f = getattr(L, fn)
            f.argtypes = [ctypes.c_float] * n
            f.restype = ctypes.c_float
        L.oracle_weno5_f64.argtypes = [ctypes.c_double] * 6
        L.oracle_weno5_f64.restype = ctypes.c_double
        L.oracle_kernel_name.restype = ctypes.c_char_p
        L.oracle_threads.restype = ctypes.c_int
        _lib = L
    return _lib


class Program:
    """Kernel names + (nops, 64) argument slots, as ctypes arrays."""

    def __init__(self, names, args):
        self.n = len(names)
        self._names = (ctypes.c_char_p * self.n)(*[n.encode() for n in names])
        self._args = np.ascontiguousarray(args, dtype=np.uint32)

    def run(self, M, first=0, last=None):
        last = self.n - 1 if last is None else last
        err = ctypes.create_string_buffer(512)
        rc = lib().oracle_run(M.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), self._names,
                              self._args.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), self.n, first, last, err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())

    def run_f64(self, M, V, first=0, last=None):
        last = self.n - 1 if last is None else last
        err = ctypes.create_string_buffer(512)
        rc = lib().oracle_run_f64(M.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                  V.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self._names,
                                  self._args.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), self.n, first, last,
                                  err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())


def region_hash(M, off, n):
    return lib().oracle_region_hash(M[off:].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)


def fnv1a(words):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    return lib().oracle_fnv1a(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(w))


def kernel_names():
    L = lib()
    return [L.oracle_kernel_name(i).decode() for i in range(L.oracle_kernel_count())]
