"""Long-horizon transposition check of the adjoint (report, CPU).

The tangent-linear build of the oracle (FP64 tangents) runs along the same
binary32 trajectory as the reverse sweep: before every dispatch its values
are reset to the oracle's float state, so both sides differentiate the same
composition of dispatches and

    <ybar, TL xdot> = <J^T ybar, xdot>

must hold to rounding however long the window, whatever the model's own
differentiability (upwind switches, WENO weights).  The reverse sweep is the
host build of the library's adjoint bodies (tests/adjoint_host.cpp), the
same bodies te_bwd runs on the GPU (tests/test_gpu_adjoint.py checks the two
sweeps agree).  Random xdot and ybar over the whole FieldArena.

    python tests/adjoint_tl_check.py [--steps 60 --nx 10 --ny 8 --nz 8]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.join(os.path.dirname(HERE), "thunderengine-web_amd")]

import adjoint_lib as al  # noqa: E402
import runs  # noqa: E402
import te_host  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--nx", type=int, default=10)
    ap.add_argument("--ny", type=int, default=8)
    ap.add_argument("--nz", type=int, default=8)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--report", type=int, default=10, help="print the dot products every this many steps")
    a = ap.parse_args()
    kw = dict(nx=a.nx, ny=a.ny, nz=a.nz, dz=20000.0 / a.nz)
    R = runs.OracleRun(te_host.config(fuse=0, **kw))
    e = R.enc
    a0, aend = e.field_offset(e.fields[0]), runs.arena_end(e)
    rng = np.random.default_rng(a.seed)
    xdot = np.zeros(aend)
    xdot[a0:] = rng.standard_normal(aend - a0)
    ybar = np.zeros(aend)
    ybar[a0:] = rng.standard_normal(aend - a0)
    for f in ("cA", "cB"):  # elimination scratch of the column solve: its final value is never read
        o = e.field_offset(f)
        ybar[o: o + e.field_words] = 0
    t0 = time.time()
    Mw = R._buf.copy()[runs.FRONT:]
    M32 = Mw.copy()
    V = al.tl_state(M32, xdot)
    checkpoints = {}
    for s in range(1, a.steps + 1):
        for i in range(R.prog.n):
            V[:, 0] = M32.view(np.float32)
            al.run_tl(R.prog, M32.copy(), V, i, i)
            R.prog.run(M32, i, i)
        if s % a.report == 0 or s == a.steps:
            checkpoints[s] = (M32.copy(), float(np.dot(ybar, V[:aend, 1])))
    print(json.dumps({"grid": "%dx%dx%d" % (e.nx, e.ny, e.nz), "steps": a.steps, "tl_s": round(time.time() - t0, 1)}),
          flush=True)
    for s, (_, lhs) in sorted(checkpoints.items()):
        Mb = np.zeros(R._buf.size, np.float32)[runs.FRONT:]
        Mb[:aend] = ybar
        Mw = R._buf.copy()[runs.FRONT:]
        al.reverse_sweep(R.names, R.args, Mw, Mb, e.field_words, steps=s)
        rhs = float(np.dot(Mb[:aend].astype(np.float64), xdot))
        print(json.dumps({"steps": s, "ybar_TL_xdot": lhs, "adjoint_ybar_xdot": rhs,
                          "rel": abs(lhs - rhs) / max(abs(lhs), abs(rhs)), "elapsed_s": round(time.time() - t0, 1)}),
              flush=True)


if __name__ == "__main__":
    main()
