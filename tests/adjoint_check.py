"""Gradient check of the adjoint at the adjoint workload's size (report).

J = sum of w^2 over 0-8 km in the 20 x 20 km column block around the bubble
after `steps` large steps of the 128 x 128 x 60 case; the gradient with
respect to the initial theta', u, v (te_adjoint on the GPU) along random unit
directions, against central differences of the oracle's FP64 build and its
FP64 tangent-linear build.  --spin runs that many steps first (from rest every
vertical upwind switch sits exactly on w = 0, where the step is not
differentiable and central differences average the two one-sided slopes).

    python tests/adjoint_check.py [--steps 60 --every 10 --spin 0 --dirs 2 --eps 1e-4]
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

WORKLOAD = dict(nx=128, ny=128, nz=60, dz=20000.0 / 60, dt=6.0, nsound=6)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--spin", type=int, default=0)
    ap.add_argument("--dirs", type=int, default=2)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--nx", type=int, default=WORKLOAD["nx"])
    ap.add_argument("--nz", type=int, default=WORKLOAD["nz"])
    a = ap.parse_args()
    kw = dict(WORKLOAD, nx=a.nx, ny=a.nx, nz=a.nz, dz=20000.0 / a.nz)
    t0 = time.time()
    run = te_host.Run(te_host.config(**kw))
    run.init_state()
    if a.spin:
        run.step(a.spin)
    R = runs.OracleRun(te_host.config(fuse=0, **kw))
    e = R.enc
    aend = runs.arena_end(e)
    for _ in range(a.spin):
        R.prog.run(R.M)
    mask = runs.state_mask(e)
    assert np.array_equal(run.read(0, aend)[mask], R.M[:aend][mask])
    c = e.nx // 2
    top = int(8000.0 / kw["dz"])
    box = te_host.Box(e, i=(c - 10, c + 10), j=(c - 10, c + 10), k=(1, top + 1))
    ctl = te_host.Controls(e, ("th", "u", "v"))
    adj = te_host.Adjoint(run, every=a.every)
    adj.forward(a.steps)
    state = run.read(0, aend).view(np.float32)
    seed = np.zeros(aend, np.float32)
    seed[box.words] = 2.0 * state[box.words]
    adj.upload(0, seed)
    adj.backward()
    grad = ctl.pull(adj.read(0, aend))
    fwd_ms, bwd_ms = adj.last_ms()
    adj.close()
    run.close()
    print(json.dumps({"grid": "%dx%dx%d" % (e.nx, e.ny, e.nz), "steps": a.steps, "every": a.every, "spin": a.spin,
                      "J": float(np.sum(state[box.words].astype(np.float64) ** 2)),
                      "fwd_ckpt_ms": round(fwd_ms, 1), "bwd_ms": round(bwd_ms, 1),
                      "gpu_done_s": round(time.time() - t0, 1)}), flush=True)
    M0 = R._buf.copy()
    rng = np.random.default_rng(41)
    for t in range(a.dirs):
        d = rng.standard_normal(ctl.size)
        d /= np.linalg.norm(d)
        Js = []
        for sgn in (1.0, -1.0):
            Mw0 = M0.copy()[runs.FRONT:]
            V = Mw0.view(np.float32).astype(np.float64)
            V[:aend] += sgn * a.eps * ctl.push(d, aend)
            for _ in range(a.steps):
                R.prog.run_f64(Mw0, V)
            Js.append(float(np.sum(V[box.words] ** 2)))
            print(json.dumps({"direction": t, "fd_leg": sgn, "elapsed_s": round(time.time() - t0, 1)}), flush=True)
        fd = (Js[0] - Js[1]) / (2 * a.eps)
        Mw0 = M0.copy()[runs.FRONT:]
        V = al.tl_state(Mw0, ctl.push(d, aend))
        for _ in range(a.steps):
            al.run_tl(R.prog, Mw0, V, 0, R.prog.n - 1)
        tl = float(np.sum(2.0 * V[box.words, 0] * V[box.words, 1]))
        ad = float(np.dot(grad, d))
        print(json.dumps({"direction": t, "adjoint": ad, "fd": fd, "tl": tl,
                          "rel_adjoint_fd": abs(ad - fd) / abs(fd), "rel_adjoint_tl": abs(ad - tl) / abs(tl),
                          "rel_tl_fd": abs(tl - fd) / abs(fd), "elapsed_s": round(time.time() - t0, 1)}), flush=True)


if __name__ == "__main__":
    main()
