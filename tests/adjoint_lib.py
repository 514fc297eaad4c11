"""Adjoint test infrastructure: the host build of the library's adjoint bodies
(tests/adjoint_host.cpp), the oracle's FP64 tangent-linear build, and a
reverse sweep over a dispatch list in Python that mirrors the library's
(save the regions each dispatch writes, run it; then restore and run its
adjoint, last dispatch first)."""
import ctypes
import os
import subprocess

import numpy as np

import oracle_lib as ol
import te_host

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "adjoint_host.cpp")
HDRS = [os.path.join(ROOT, "thunderengine-web_amd", "csrc", h) for h in ("adj_core.hpp", "adj_bodies.hpp")]
OUT = os.path.join(HERE, "_build", "libadjhost.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        stale = not os.path.exists(OUT) or any(os.path.getmtime(f) > os.path.getmtime(OUT) for f in [SRC] + HDRS)
        if stale:
            os.makedirs(os.path.dirname(OUT), exist_ok=True)
            subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-mfma", "-ffp-contract=off",
                                   "-fno-fast-math", "-o", OUT, SRC])
        L = ctypes.CDLL(OUT)
        L.adjhost_op.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_int]
        L.adjhost_writes.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.adjhost_name.restype = ctypes.c_char_p
        _lib = L
    return _lib


def kernel_names():
    L = lib()
    return [L.adjhost_name(i).decode() for i in range(L.adjhost_count())]


def writes(name, A, field_words):
    """(offset, words) regions dispatch (name, A) writes."""
    if name == "_copy":
        return [(int(A[5]), int(A[0]))]
    out = (ctypes.c_int * 8)()
    n = lib().adjhost_writes(name.encode(), out)
    if n < 0:
        raise KeyError("no adjoint for " + name)
    regs = []
    for q in out[:n]:
        off = int(A[4 + q])
        if all(off != r[0] for r in regs):
            regs.append((off, field_words))
    return regs


def host_adjoint(name, A, Mw, Mb, scatter_only=False):
    """Adds dispatch (name, A)'s J^T to adjoint memory Mb (float32, same indexing
    as the uint32 word memory Mw); scatter_only: the per-point body's scatter
    form even where the library runs a gather-form adjoint."""
    A = np.ascontiguousarray(A, dtype=np.uint32)
    rc = lib().adjhost_op(name.encode(), A.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                          Mw.ctypes.data, Mb.ctypes.data, 1 if scatter_only else 0)
    if rc != 0:
        raise KeyError("no adjoint for " + name)


def run_tl(prog, Mw, V, first, last):
    """Oracle FP64 tangent-linear run of ops [first, last]; V is (words, 2) float64
    (value, tangent), Mw the uint32 words (parameters)."""
    err = ctypes.create_string_buffer(512)
    L = ol.lib()
    if not hasattr(L, "_tl_ready"):
        L.oracle_run_tl.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_uint32), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L._tl_ready = True
    rc = L.oracle_run_tl(Mw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                         V.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), prog._names,
                         prog._args.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), prog.n, first, last, err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())


def tl_state(Mw, tangent):
    """(value, tangent) pairs for the TL build: values are the float words as
    doubles (parameter words are read from Mw by the kernels)."""
    V = np.zeros((Mw.size, 2), dtype=np.float64)
    V[:, 0] = Mw.view(np.float32)
    V[: tangent.size, 1] = tangent
    return V


def reverse_sweep(names, args, Mw, Mb, field_words, first=0, last=None, steps=1):
    """Reverse of `steps` repetitions of ops [first, last] from state Mw (uint32
    words, modified: ends at the state after the forward run) with Mb the
    adjoint of the final state (float32, updated in place to the adjoint of
    the initial state).  Forward by the oracle, adjoints by the host build."""
    last = len(names) - 1 if last is None else last
    prog = ol.Program(names, args)
    logs = []
    for _ in range(steps):
        for i in range(first, last + 1):
            saved = [(o, n, Mw[o: o + n].copy()) for (o, n) in writes(names[i], args[i], field_words)]
            logs.append((i, saved))
            prog.run(Mw, i, i)
    for i, saved in reversed(logs):
        for o, n, w in saved:
            Mw[o: o + n] = w
        host_adjoint(names[i], args[i], Mw, Mb)


# the control mapping is part of the host API
Box = te_host.Box
Controls = te_host.Controls
