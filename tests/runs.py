"""Programs of the encoder run in the oracle (test infrastructure).

A configuration is encoded by libte.so without a device (te_encode), its t=0
state filled into the host image (te_init_state), and the same program and
memory image are then run by the oracle; GPU tests run them through te_create
and compare word for word.
"""
import numpy as np

import oracle_lib as ol
import te_host

FRONT = 1 << 16  # words below M that a stencil at a region's start may touch


def unfused(cfg):
    """Copy of a te_config that encodes the reference's dispatch list one for one."""
    c = type(cfg).from_buffer_copy(cfg)
    c._keep = getattr(cfg, "_keep", None)
    c.fuse = 0
    return c


class OracleRun:
    """The oracle stepping the reference's dispatch list for `cfg` (fuse=0 unless
    fused=True, whose merged launches the oracle runs as the dispatches they
    replace)."""

    def __init__(self, cfg, fused=False):
        if not fused:
            cfg = unfused(cfg)
        self.enc = te_host.Run(cfg, encode_only=True)
        self.enc.init_state()
        e = self.enc
        self.names, self.args = e.program()
        self.prog = ol.Program(self.names, self.args)
        self.words = e.mem_words
        back = e.sJ * e.sK * (e.nz + 4) + 4096
        self._buf = np.zeros(FRONT + self.words + back, dtype=np.uint32)
        self.M = self._buf[FRONT:]
        self.M[: self.words] = e.read(0, self.words)
        self.M0 = self.M[: self.words].copy()

    def step(self, n=1):
        for _ in range(n):
            self.prog.run(self.M)

    def field(self, name):
        e = self.enc
        o = e.field_offset(name)
        return self.M[o: o + e.sJ * e.sK * (e.nz + 1)].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)

    def interior(self, name):
        e = self.enc
        return self.field(name)[: e.nz, 3: 3 + e.ny, 3: 3 + e.nx]


# Scratch fields the fused launches may leave with other values than the
# reference's dispatch list (fluxes, elimination coefficients, fpk, tend,
# sadv, and the acoustic substep's per-stage terms qk, mm, pk1, pk2, which
# the one-launch substep forms per level instead of storing).
# tests/test_oracle_invariants.py checks that each is dead at the start of a
# step: its contents then never reach another field.
FUSED_SCRATCH = ("fx", "fy", "fz", "cA", "cB", "fpk", "tend", "sadv", "tketen_s", "qk", "mm", "pk1", "pk2")


def state_mask(e):
    """Boolean mask over [0, arena_end) of the words compared in fused runs:
    profiles, parameter regions and every arena field but FUSED_SCRATCH."""
    end = arena_end(e)
    m = np.ones(end, dtype=bool)
    for f in FUSED_SCRATCH:
        if f in e.fields:
            o = e.field_offset(f)
            m[o: o + e.field_words] = False
    return m


def arena_end(e):
    """Word offset just past the FieldArena (profiles and arena lie below it)."""
    return e.field_offset(e.fields[0]) + e.field_words * len(e.fields)


def first_difference(run, a, b):
    """Names the first arena field where word arrays a and b differ."""
    e = run.enc if hasattr(run, "enc") else run
    diff = np.nonzero(a != b)[0]
    if diff.size == 0:
        return None
    w = int(diff[0])
    a0 = e.field_offset(e.fields[0])
    if w >= a0:
        q = (w - a0) // e.field_words
        if q < len(e.fields):
            r = w - a0 - q * e.field_words
            k, rem = divmod(r, e.sJ * e.sK)
            j, i = divmod(rem, e.sJ)
            return "%s[k=%d j=%d i=%d] (%d words differ)" % (e.fields[q], k, j - 3, i - 3, diff.size)
    return "word %d (%d words differ)" % (w, diff.size)
