"""Variable large step (adapt_dt = 1) on the MI355X against the oracle.

The GPU run (te_adaptive_start / te_adaptive_step: stability_max at the end
of each step, the two words read and cleared, the clock, te_set_time_step
with graph reuse while nsound stays) must take the oracle's steps: the same
dt and nsound sequence, landing on the 60 s output times, and the state bit
for bit after every step.  The sequence crosses nsound 6 -> 8 -> 6, so both
the parameter-only update and the replaced dispatch list run.
"""
import sys

import numpy as np
import pytest

import runs
import te_host
from adaptive_lib import OracleAdaptive

pytestmark = pytest.mark.gpu

CFG = dict(nx=32, ny=28, nz=16, dx=1000.0, dy=1000.0, dz=500.0, dt=6.0, tke=1, initial_tke=3, rayleigh=1,
           adaptive_dt=1, tapfrq=60.0, bubble_amp=3.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("fuse", [1, 0])
def test_adaptive_run_matches_the_oracle(gpu, fuse):
    cfg = te_host.config(fuse=fuse, **CFG)
    say("adaptive 32x28x16 TKE, fuse", fuse)
    run = te_host.Run(cfg)
    run.init_state()
    ref = OracleAdaptive(cfg)
    end = runs.arena_end(run)
    keep = runs.state_mask(run) if fuse else np.ones(end, dtype=bool)

    # the t = 0 maxima: stability_max alone, words compared before the clock reads them
    names, _ = run.program()
    i = len(names) - 1 - names[::-1].index("stability_max")
    run.run(1, i, i)
    run.sync()
    so = run.region_offset("stability")
    want = ref.stability_words()
    got = run.read(so, 2)
    assert np.array_equal(got, want), (got, want)
    run.upload(so, np.zeros(2, dtype=np.uint32))

    run.adaptive_start()
    ref.start(1000.0)
    seen = set()
    for s in range(14):
        g, o = run.clock(), ref.clock.state()
        assert (g["dt"], g["nsound"], g["mtime"]) == (o["dt"], o["nsound"], o["mtime"]), (s, g, o)
        seen.add(g["nsound"])
        run.adaptive_step(1)
        ref.step()
        a, b = run.read(0, end), ref.M[:end].copy()
        a[~keep] = 0
        b[~keep] = 0
        assert np.array_equal(a, b), (s, runs.first_difference(run, a, b))
        say("step", s + 1, "dt", g["dt"], "nsound", g["nsound"], "bit-exact")
    st = run.clock()
    assert st["mtime"] > 120.0 and {6, 8} <= seen
    run.close()
