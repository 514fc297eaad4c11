"""The oracle's adaptive run (adapt_dt = 1) for parity tests: the oracle
steps the reference's dispatch list, the clock (te_clock) chooses dt and
nsound from the oracle's own stability maxima, and te_set_time_step on an
encode-only context gives each step's program and parameter words
(web/core/driver.js:36-50, web/view.js:513-550)."""
import numpy as np

import oracle_lib as ol
import runs
import te_host


class OracleAdaptive:
    def __init__(self, cfg):
        self.enc = te_host.Run(runs.unfused(cfg), encode_only=True)
        e = self.enc
        e.init_state()
        self.words = e.mem_words
        self._buf = np.zeros(runs.FRONT + self.words + e.sJ * e.sK * (e.nz + 4) + 4096, dtype=np.uint32)
        self.M = self._buf[runs.FRONT: runs.FRONT + self.words]
        self.M[:] = e.read(0, self.words)
        self.so = e.region_offset("stability")
        self.pay0 = runs.arena_end(e)
        self.cfg = cfg
        self._program()

    def _program(self):
        self.names, self.args = self.enc.program()
        self.prog = ol.Program(self.names, self.args)
        self.M[self.pay0:] = self.enc.read(self.pay0, self.words - self.pay0)

    def stability_words(self):
        """Runs the step's stability_max dispatch alone; returns and clears the two words."""
        i = len(self.names) - 1 - self.names[::-1].index("stability_max")
        self.prog.run(self._buf[runs.FRONT:], i, i)
        w = self.M[self.so: self.so + 2].copy()
        self.M[self.so: self.so + 2] = 0
        return w

    @staticmethod
    def maxima(w):
        v = w.view(np.float32).astype(np.float64)
        return float(np.float32(np.sqrt(v[0]))), float(np.float32(np.sqrt(v[1])))

    def start(self, min_dx):
        c = self.cfg
        self.clock = te_host.Clock(c.dt, min_dx, freq={"tape": c.tapfrq, "restart": c.rstfrq, "stat": c.statfrq})
        self.clock.start(*self.maxima(self.stability_words()))
        self._retime()

    def _retime(self):
        st = self.clock.state()
        self.enc.set_time_step(st["dt"], st["nsound"])
        self._program()

    def step(self):
        self.clock.begin_step()
        self.prog.run(self._buf[runs.FRONT:])
        w = self.M[self.so: self.so + 2].copy()
        self.M[self.so: self.so + 2] = 0
        self.clock.end_step(*self.maxima(w))
        self._retime()
