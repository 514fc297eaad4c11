"""Adjoint of the large step, checked without a GPU.

The library's adjoint bodies (thunderengine-web_amd/csrc/adj_bodies.hpp) are
built for the host (tests/adjoint_host.cpp) and checked against the oracle's
FP64 tangent-linear build (oracle.hpp, T = TL):
  * per dispatch: <ybar, J xdot> = <J^T ybar, xdot> for random xdot, ybar over
    the FieldArena, with the dispatch's input state taken from a real step;
  * per step: the reverse sweep (save written regions, run; then restore and
    run the adjoints, last dispatch first) over whole steps, the same identity;
  * against FP64 central differences of J = sum of w^2 in a box (the oracle's
    double build), along random unit directions of the initial theta', u, v, qv.
The GPU tests run the same checks through the C ABI (te_adj_ops, te_bwd).
"""
import numpy as np
import pytest

import adjoint_lib as al
import runs
import te_host

SMALL = dict(nx=10, ny=8, nz=8, dz=2500.0)
CONFIGS = {
    "moist_weno_rayleigh": dict(SMALL),
    "dry_weno_momentum": dict(SMALL, moist=0, weno_momentum=1),
}
DEAD = ("cA", "cB")  # elimination scratch of the column solve: never read after it
GATHERED = ("adv_assemble", "acoustic_uv", "acoustic_fpk", "acoustic_pp", "adv_flux_x", "adv_flux_y", "adv_flux_x_stag",
            "adv_flux_y_stag", "acoustic_accumulate")


def _arena(e):
    return e.field_offset(e.fields[0]), runs.arena_end(e)


def _zero_fields(e, vec, names):
    for f in names:
        if f in e.fields:
            o = e.field_offset(f)
            vec[o: o + e.field_words] = 0


def test_every_dispatch_of_the_path_has_an_adjoint():
    have = set(al.kernel_names())
    for kw in list(CONFIGS.values()) + [dict(SMALL, rayleigh=0, weno_scalars=0)]:
        r = te_host.Run(te_host.config(fuse=0, **kw), encode_only=True)
        names, _ = r.program()
        missing = sorted(set(names) - have)
        assert not missing, (kw, missing)
        r.close()


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_dispatch_adjoints_transpose_the_tangent_linear_build(name):
    R = runs.OracleRun(te_host.config(fuse=0, **CONFIGS[name]))
    e = R.enc
    a0, aend = _arena(e)
    rng = np.random.default_rng(7)
    worst = {}
    for i, nm in enumerate(R.names):
        Mw = R._buf.copy()[runs.FRONT:]
        xdot = np.zeros(aend)
        xdot[a0:] = rng.standard_normal(aend - a0)
        ybar = np.zeros(aend)
        ybar[a0:] = rng.standard_normal(aend - a0)
        if nm == "acoustic_w_solve":
            _zero_fields(e, ybar, DEAD)
        V = al.tl_state(Mw, xdot)
        al.run_tl(R.prog, Mw.copy(), V, i, i)
        Mb = np.zeros(R._buf.size, np.float32)[runs.FRONT:]
        Mb[:aend] = ybar
        al.host_adjoint(nm, R.args[i], Mw, Mb)
        lhs = float(np.dot(ybar, V[:aend, 1]))
        rhs = float(np.dot(Mb[:aend].astype(np.float64), xdot))
        worst[nm] = max(worst.get(nm, 0.0), abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300))
        R.prog.run(R.M, i, i)
    bad = {k: v for k, v in worst.items() if v > 1e-4}
    assert not bad, bad
    assert len(worst) >= 30


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_gather_adjoints_equal_the_scatter_form(name):
    """The kernels in GATHERED run gather-form adjoints in the library; word for word they equal the per-point bodies' scatter form
    (up to binary32 against binary64 rounding of the Jacobian entries)."""
    R = runs.OracleRun(te_host.config(fuse=0, **CONFIGS[name]))
    e = R.enc
    a0, aend = _arena(e)
    rng = np.random.default_rng(13)
    seen = set()
    for i, nm in enumerate(R.names):
        if nm in GATHERED:
            Mw = R._buf.copy()[runs.FRONT:]
            ybar = np.zeros(aend, np.float32)
            ybar[a0:] = rng.standard_normal(aend - a0)
            out = []
            for scatter in (False, True):
                Mb = np.zeros(R._buf.size, np.float32)[runs.FRONT:]
                Mb[:aend] = ybar
                al.host_adjoint(nm, R.args[i], Mw, Mb, scatter_only=scatter)
                out.append(Mb[:aend].astype(np.float64))
            err = np.abs(out[0] - out[1]).max() / np.abs(out[1]).max()
            assert err < 1e-5, (nm, i, err)
            seen.add(nm)
        R.prog.run(R.M, i, i)
    assert seen == set(GATHERED)


def test_reverse_sweep_transposes_two_steps():
    """The tangent-linear build runs along the same binary32 trajectory as the
    sweep (its values reset to the oracle's float state before each
    dispatch), so both sides differentiate the same composition."""
    R = runs.OracleRun(te_host.config(fuse=0, **SMALL))
    e = R.enc
    a0, aend = _arena(e)
    rng = np.random.default_rng(11)
    xdot = np.zeros(aend)
    xdot[a0:] = rng.standard_normal(aend - a0)
    ybar = np.zeros(aend)
    ybar[a0:] = rng.standard_normal(aend - a0)
    _zero_fields(e, ybar, DEAD)
    Mw = R._buf.copy()[runs.FRONT:]
    M0 = Mw.copy()
    M32 = Mw.copy()
    V = al.tl_state(M32, xdot)
    for _ in range(2):
        for i in range(R.prog.n):
            V[:, 0] = M32.view(np.float32)
            al.run_tl(R.prog, M32.copy(), V, i, i)
            R.prog.run(M32, i, i)
    Mb = np.zeros(R._buf.size, np.float32)[runs.FRONT:]
    Mb[:aend] = ybar
    al.reverse_sweep(R.names, R.args, Mw, Mb, e.field_words, steps=2)
    assert np.array_equal(Mw, M0)  # the restores undo every write of the two steps
    lhs = float(np.dot(ybar, V[:aend, 1]))
    rhs = float(np.dot(Mb[:aend].astype(np.float64), xdot))
    assert abs(lhs - rhs) <= 1e-4 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def _spun_up(kw, spin):
    """Oracle run after `spin` forward steps: away from the state of rest,
    where every vertical upwind switch sits exactly on w = 0."""
    R = runs.OracleRun(te_host.config(fuse=0, **kw))
    for _ in range(spin):
        R.prog.run(R.M)
    return R


def _adjoint_gradient(R, box, ctl, steps):
    e = R.enc
    M0 = R._buf.copy()
    Mw = M0.copy()[runs.FRONT:]
    for _ in range(steps):
        R.prog.run(Mw)
    Mb = np.zeros(R._buf.size, np.float32)[runs.FRONT:]
    Mb[box.words] = 2.0 * Mw[box.words].view(np.float32)
    Mw = M0.copy()[runs.FRONT:]
    al.reverse_sweep(R.names, R.args, Mw, Mb, e.field_words, steps=steps)
    return ctl.pull(Mb)


def test_gradient_matches_fp64_central_differences_and_tangent_linear():
    """J = sum of w^2 over levels 1..4 of the central 4 x 4 columns after two
    steps from a spun-up state (3 steps after t=0); its gradient with respect
    to theta', u, v (each perturbation moves a field, its _a copy and their
    periodic halo images) along 3 random unit directions, against
    (J(x + eps d) - J(x - eps d)) / (2 eps) of the FP64 build and against the
    FP64 tangent-linear build.  eps = 1e-4: at eps = 1e-3 the differences'
    own truncation error reaches 5e-4 relative on this grid (they converge to
    the tangent-linear value as eps shrinks).  The adjoint linearises along
    the binary32 trajectory; measured agreement is 4e-6..1.2e-4 relative."""
    R = _spun_up(SMALL, 3)
    e = R.enc
    aend = runs.arena_end(e)
    box = al.Box(e, i=(3, 7), j=(2, 6), k=(1, 5))
    ctl = al.Controls(e, ("th", "u", "v"))
    grad = _adjoint_gradient(R, box, ctl, 2)
    M0 = R._buf.copy()
    rng = np.random.default_rng(3)
    for _ in range(3):
        d = rng.standard_normal(ctl.size)
        d /= np.linalg.norm(d)
        ad = float(np.dot(grad, d))
        Js = []
        for sgn in (1.0, -1.0):
            Mw0 = M0.copy()[runs.FRONT:]
            V = Mw0.view(np.float32).astype(np.float64)
            V[:aend] += sgn * 1e-4 * ctl.push(d, aend)
            R.prog.run_f64(Mw0, V)
            R.prog.run_f64(Mw0, V)
            Js.append(float(np.sum(V[box.words] ** 2)))
        fd = (Js[0] - Js[1]) / 2e-4
        Mw0 = M0.copy()[runs.FRONT:]
        V = al.tl_state(Mw0, ctl.push(d, aend))
        for _ in range(2):
            al.run_tl(R.prog, Mw0, V, 0, R.prog.n - 1)
        tl = float(np.sum(2.0 * V[box.words, 0] * V[box.words, 1]))
        assert abs(ad - fd) <= 3e-4 * abs(fd), (ad, fd)
        assert abs(ad - tl) <= 3e-4 * abs(tl), (ad, tl)


def test_moisture_gradient_matches_the_tangent_linear_build():
    """Along directions that include qv the FP64 model is not differentiable
    where qv is horizontally uniform: WENO-Z's weights (weps = 1e-20) jump from
    the ideal weights to data-driven ones for any perturbation above ~1e-10,
    so central differences do not converge to a derivative there.  The adjoint
    is checked against the tangent-linear build (the derivative of the branch
    the step took) instead."""
    R = _spun_up(SMALL, 3)
    e = R.enc
    aend = runs.arena_end(e)
    box = al.Box(e, i=(3, 7), j=(2, 6), k=(1, 5))
    ctl = al.Controls(e, ("th", "u", "v", "qv"))
    grad = _adjoint_gradient(R, box, ctl, 2)
    M0 = R._buf.copy()
    rng = np.random.default_rng(5)
    for _ in range(2):
        d = rng.standard_normal(ctl.size)
        d /= np.linalg.norm(d)
        Mw0 = M0.copy()[runs.FRONT:]
        V = al.tl_state(Mw0, ctl.push(d, aend))
        for _ in range(2):
            al.run_tl(R.prog, Mw0, V, 0, R.prog.n - 1)
        tl = float(np.sum(2.0 * V[box.words, 0] * V[box.words, 1]))
        ad = float(np.dot(grad, d))
        assert abs(ad - tl) <= 1e-3 * abs(tl), (ad, tl)
