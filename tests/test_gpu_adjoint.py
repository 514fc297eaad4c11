"""Adjoint of the large step on the MI355X, through the C ABI (te_adjoint_*).

* every dispatch's device adjoint (te_adj_ops) against the oracle's FP64
  tangent-linear build: <ybar, J xdot> = <J^T ybar, xdot>;
* the library's reverse sweep (te_fwd_ckpt on the fused fast path, te_bwd on
  the unfused dispatch list with checkpoints every 1 and every 2 steps)
  against the host build's reverse sweep over the oracle's float steps;
* the gradient of J = sum of w^2 over a box from a spun-up state against FP64
  central differences and the FP64 tangent-linear build.
"""
import sys

import numpy as np
import pytest

import adjoint_lib as al
import runs
import te_host

pytestmark = pytest.mark.gpu

GRID = dict(nx=20, ny=16, nz=10, dz=2000.0)
DEAD = ("cA", "cB")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _zero_fields(e, vec, names):
    for f in names:
        if f in e.fields:
            o = e.field_offset(f)
            vec[o: o + e.field_words] = 0


@pytest.mark.parametrize("kw", [dict(GRID), dict(GRID, moist=0, weno_momentum=1)], ids=["moist", "dry_weno_uv"])
def test_gpu_dispatch_adjoints_transpose_the_tangent_linear_build(gpu, kw):
    cfg = te_host.config(fuse=0, **kw)  # the adjoint's program is then the reference's dispatch list
    say("dispatch adjoints", kw)
    R = runs.OracleRun(te_host.config(fuse=0, **kw))
    e = R.enc
    run = te_host.Run(cfg)
    adj = te_host.Adjoint(run, every=1)
    prog = adj.program
    names, _ = prog.program()
    assert names == R.names
    a0, aend = e.field_offset(e.fields[0]), runs.arena_end(e)
    rng = np.random.default_rng(17)
    worst = {}
    for i, nm in enumerate(R.names):
        Mw = R.M.copy()
        xdot = np.zeros(aend)
        xdot[a0:] = rng.standard_normal(aend - a0)
        ybar = np.zeros(aend)
        ybar[a0:] = rng.standard_normal(aend - a0)
        if nm == "acoustic_w_solve":
            _zero_fields(e, ybar, DEAD)
        V = al.tl_state(R._buf.copy()[runs.FRONT:], xdot)
        al.run_tl(R.prog, R._buf.copy()[runs.FRONT:], V, i, i)
        prog.upload(0, Mw[: R.words])
        adj.zero()
        adj.upload(0, ybar.astype(np.float32))
        adj.dispatch_adjoints(i, i)
        xbar = adj.read(0, aend).astype(np.float64)
        lhs = float(np.dot(ybar.astype(np.float32).astype(np.float64), V[:aend, 1]))
        rhs = float(np.dot(xbar, xdot))
        worst[nm] = max(worst.get(nm, 0.0), abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300))
        R.prog.run(R.M, i, i)
    adj.close()
    run.close()
    say("worst", sorted(worst.items(), key=lambda kv: -kv[1])[:5])
    bad = {k: v for k, v in worst.items() if v > 1e-4}
    assert not bad, bad


@pytest.mark.parametrize("every,segment_log,fuse", [(1, True, 1), (2, True, 1), (2, False, 1), (2, True, 0)])
def test_gpu_reverse_sweep_matches_the_host_sweep(gpu, every, segment_log, fuse, monkeypatch):
    """fuse 1: the adjoint recomputes the fused program and runs each fused
    launch's adjoint as the reference dispatches it stands for; fuse 0: the
    reference's list.  segment_log False: a zero log budget forces one step's
    log at a time with the segment's step-start states recomputed first."""
    steps = 3
    say("reverse sweep, steps", steps, "checkpoint every", every, "segment log", segment_log, "fuse", fuse)
    if not segment_log:
        monkeypatch.setenv("TE_ADJOINT_LOG_GIB", "0")
    run = te_host.Run(te_host.config(fuse=fuse, **GRID))
    run.init_state()
    adj = te_host.Adjoint(run, every=every)
    R = runs.OracleRun(te_host.config(fuse=0, **GRID))
    e = R.enc
    a0, aend = e.field_offset(e.fields[0]), runs.arena_end(e)
    M0 = R.M.copy()
    assert np.array_equal(run.read(0, aend), M0[:aend])
    adj.forward(steps)
    rng = np.random.default_rng(23)
    ybar = np.zeros(aend, np.float32)
    ybar[a0:] = rng.standard_normal(aend - a0)
    # the scratch fields' final values are not part of the fused step's
    # results (its launches keep fluxes and coefficients in registers)
    _zero_fields(e, ybar, runs.FUSED_SCRATCH if fuse else DEAD)
    adj.upload(0, ybar)
    adj.backward()
    dev = adj.read(0, aend).astype(np.float64)
    Mb = np.zeros(R._buf.size, np.float32)[runs.FRONT:]
    Mb[:aend] = ybar
    al.reverse_sweep(R.names, R.args, R.M, Mb, e.field_words, steps=steps)
    assert np.array_equal(R.M, M0)
    host = Mb[:aend].astype(np.float64)
    err = np.linalg.norm(dev - host) / np.linalg.norm(host)
    fwd_ms, bwd_ms = adj.last_ms()
    say("relative difference", err, "fwd ms", fwd_ms, "bwd ms", bwd_ms)
    adj.close()
    run.close()
    assert err < 1e-5


def test_gpu_gradient_matches_fp64_central_differences(gpu):
    """J = sum of w^2 over levels 1..5 of the central 6 x 6 columns after
    3 steps from the state 3 steps after t=0; gradient with respect to
    theta', u, v along 2 random unit directions against central differences
    of the FP64 build (eps = 1e-4) and the FP64 tangent-linear build."""
    spin, steps = 3, 3
    run = te_host.Run(te_host.config(**GRID))
    run.init_state()
    run.step(spin)
    adj = te_host.Adjoint(run, every=2)
    R = runs.OracleRun(te_host.config(fuse=0, **GRID))
    e = R.enc
    aend = runs.arena_end(e)
    for _ in range(spin):
        R.prog.run(R.M)
    mask = runs.state_mask(e)  # the fused launches' dead scratch fields differ
    assert np.array_equal(run.read(0, aend)[mask], R.M[:aend][mask])
    box = al.Box(e, i=(7, 13), j=(5, 11), k=(1, 6))
    ctl = al.Controls(e, ("th", "u", "v"))
    adj.forward(steps)
    wfin = run.read(0, aend).view(np.float32)
    seed = np.zeros(aend, np.float32)
    seed[box.words] = 2.0 * wfin[box.words]
    adj.upload(0, seed)
    adj.backward()
    grad = ctl.pull(adj.read(0, aend))
    adj.close()
    run.close()
    M0 = R._buf.copy()
    rng = np.random.default_rng(29)
    for t in range(2):
        d = rng.standard_normal(ctl.size)
        d /= np.linalg.norm(d)
        Js = []
        for sgn in (1.0, -1.0):
            Mw0 = M0.copy()[runs.FRONT:]
            V = Mw0.view(np.float32).astype(np.float64)
            V[:aend] += sgn * 1e-4 * ctl.push(d, aend)
            for _ in range(steps):
                R.prog.run_f64(Mw0, V)
            Js.append(float(np.sum(V[box.words] ** 2)))
        fd = (Js[0] - Js[1]) / 2e-4
        Mw0 = M0.copy()[runs.FRONT:]
        V = al.tl_state(Mw0, ctl.push(d, aend))
        for _ in range(steps):
            al.run_tl(R.prog, Mw0, V, 0, R.prog.n - 1)
        tl = float(np.sum(2.0 * V[box.words, 0] * V[box.words, 1]))
        ad = float(np.dot(grad, d))
        say("direction", t, "adjoint", ad, "fd", fd, "tl", tl)
        assert abs(ad - fd) <= 3e-4 * abs(fd), (ad, fd)
        assert abs(ad - tl) <= 3e-4 * abs(tl), (ad, tl)


def test_gpu_backward_twice_and_closed_run(gpu):
    """Two backward passes on one Adjoint (zero, seed, te_bwd each time: a
    training loop) give the same gradient to the float-atomics' reordering,
    and an Adjoint whose run was closed first raises instead of using it."""
    run = te_host.Run(te_host.config(**GRID))
    run.init_state()
    adj = te_host.Adjoint(run, every=2)
    aend = runs.arena_end(run)
    a0 = run.field_offset(run.fields[0])
    adj.forward(2)
    rng = np.random.default_rng(31)
    seed = np.zeros(aend, np.float32)
    seed[a0:] = rng.standard_normal(aend - a0)
    _zero_fields(run, seed, runs.FUSED_SCRATCH)
    grads = []
    for rep in range(2):
        say("backward pass", rep)
        adj.zero()
        adj.upload(0, seed)
        adj.backward()
        grads.append(adj.read(0, aend).astype(np.float64))
    err = np.linalg.norm(grads[0] - grads[1]) / np.linalg.norm(grads[0])
    say("relative difference between the passes", err)
    assert err < 1e-6
    run.close()
    with pytest.raises(te_host.TeError, match="destroyed"):
        adj.forward(1)
    adj.close()
