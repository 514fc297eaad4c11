"""Variable large step (adapt_dt = 1), CPU side.

The clock (te_clock_*, web/core/adaptive.js:15-109, CM1 getnewdt / dtsmall)
is pinned by what CM1's rules fix in closed form:

* Courant number below 1: dt grows by at most 10 % per step, up to 2 dtl;
  above 1: dt shrinks by the factor 1 / cfl;
* the diffusion limit caps dt at dt * 0.09 / ks;
* nsound is the smallest even number >= 4 with dt / nsound <= dtsm =
  0.6 dmin / 360, capped at 192 (then dt = 192 dtsm);
* output clocks: the steps land on the output times (tout split in 2 or 3
  equal steps ahead of it), and the clock fires at those steps.

The stability maxima (stability_max) are pinned on uniform fields, whose
Courant and diffusion numbers are closed forms, and te_set_time_step's
re-encoding by its substep count, its parameter bytes and a round trip.
"""
import math

import numpy as np
import pytest

import oracle_lib as ol
import runs
import te_host

f32 = np.float32


def dtsm(dmin):
    return float(f32(f32(f32(0.6) * f32(dmin)) / f32(360.0)))


def smallest_even(dbldt, dmin):
    n = 4
    while dbldt / n > dtsm(dmin):
        n += 2
    return n


def test_growth_is_ten_percent_per_step_up_to_twice_dtl():
    c = te_host.Clock(6.0, 1000.0, freq={"tape": 0.0, "restart": 0.0, "stat": 0.0})
    c.start(0.5, 0.0)
    want = 6.0
    for step in range(12):
        want = min(want * float(f32(1.1)), 12.0)
        st = c.state()
        assert st["dbldt"] == pytest.approx(want, rel=1e-12), step
        assert st["dt"] == f32(st["dbldt"])
        assert st["nsound"] == smallest_even(st["dbldt"], 1000.0)
        c.begin_step()
        c.end_step(0.5, 0.0)
    assert c.state()["dbldt"] == 12.0


def test_courant_above_one_shrinks_by_its_inverse():
    c = te_host.Clock(10.0, 2000.0)
    c.start(2.0, 0.0)
    assert c.state()["dbldt"] == pytest.approx(5.0, rel=1e-7)
    c.begin_step()
    c.end_step(1.25, 0.0)
    assert c.state()["dbldt"] == pytest.approx(5.0 * float(f32(1 / f32(1.25))), rel=1e-12)
    # the reported maxima are rescaled to the new dt (getnewdt's tem)
    st = c.state()
    assert st["cflmax"] == pytest.approx(1.25 * st["dt"] / st["dtlast"], rel=1e-6)


def test_diffusion_limit():
    c = te_host.Clock(6.0, 1000.0)
    c.start(0.1, 0.0)  # dt 6.6
    dt0 = c.state()["dt"]
    c.begin_step()
    c.end_step(0.1, 0.5)  # ks * dt' / dt = 0.5 * 1.1 > 0.09: dt' = dt 0.09 / 0.5
    assert c.state()["dbldt"] == pytest.approx(dt0 * 0.09 / 0.5, rel=1e-6)


@pytest.mark.parametrize("seed", range(4))
def test_nsound_is_the_smallest_even_count(seed):
    rng = np.random.default_rng(seed)
    for dmin in (100.0, 500.0, 1000.0, 3000.0):
        for cfl in rng.uniform(0.05, 3.0, size=20):
            c = te_host.Clock(float(rng.uniform(0.5, 20.0)), dmin)
            c.start(float(cfl), 0.0)
            st = c.state()
            assert st["nsound"] % 2 == 0 and st["nsound"] >= 4
            assert st["nsound"] == min(smallest_even(st["dbldt"], dmin), 192), (dmin, cfl, st)


def test_nsound_cap():
    c = te_host.Clock(6.0, 1.0)  # dtsm = 1/600 s: 3960 substeps wanted
    c.start(0.5, 0.0)
    st = c.state()
    assert st["nsound"] == 192
    assert st["dt"] == f32(f32(192.0) * f32(dtsm(1.0)))


def test_steps_land_on_output_times():
    c = te_host.Clock(7.0, 1000.0, freq={"tape": 60.0, "restart": 0.0, "stat": 0.0})
    c.start(0.2, 0.0)
    hits = []
    for _ in range(60):
        c.begin_step()
        st = c.state()
        if "tape" in c.state()["fired_kinds"]:
            hits.append(st["mtime"])
        c.end_step(0.2, 0.0)
        assert c.state()["dbldt"] <= 14.0
    assert len(hits) >= 5
    for t in hits:
        assert abs(t / 60.0 - round(t / 60.0)) < 1e-9 and round(t / 60.0) >= 1, hits
    assert [round(t / 60.0) for t in hits] == list(range(1, len(hits) + 1))


# ---- stability maxima and the re-encoded step

ADAPT = dict(nx=24, ny=20, nz=12, dz=500.0, dx=1000.0, dy=800.0, tke=1, initial_tke=3, adaptive_dt=1, dt=6.0,
             fuse=0, tapfrq=0.0, statfrq=0.0, rstfrq=0.0)


def stability_op(run):
    names, A = run.program()
    i = len(names) - 1 - names[::-1].index("stability_max")
    return names, A, i


def test_stability_max_on_uniform_fields():
    """u = U, v = V, w = W, kmh = khh = kmv = K everywhere: sqrt(c2) =
    dt sqrt((U/dx)^2 + (V/dy)^2 + (W/dz)^2), ks = dt K sqrt(1/dx^4 + 1/dy^4);
    one larger u at a point raises cflmax to that point's value."""
    cfg = te_host.config(**ADAPT)
    e = te_host.Run(cfg, encode_only=True)
    names, A, i = stability_op(e)
    assert names[i + 1:] and all(n == "_copy" for n in names[i + 1:])  # before the state swap
    M = e.read(0, e.mem_words).copy()
    U, V, W, K = 12.0, -7.0, 3.0, 150.0
    for name, val in (("u", U), ("v", V), ("w", W), ("kmh", K), ("khh", K), ("kmv", K)):
        o = e.field_offset(name)
        M[o: o + e.field_words] = np.full(e.field_words, val, dtype=np.float32).view(np.uint32)
    so = e.region_offset("stability")
    M[so: so + 2] = 0
    buf = np.zeros(runs.FRONT + M.size + e.sJ * e.sK * (e.nz + 4), dtype=np.uint32)
    buf[runs.FRONT: runs.FRONT + M.size] = M
    view = buf[runs.FRONT:]
    ol.Program([names[i]], A[i:i + 1]).run(view)
    cfl, ks = np.sqrt(view[so: so + 2].view(np.float32).astype(np.float64))
    dt = 6.0
    assert cfl == pytest.approx(dt * math.sqrt((U / 1000) ** 2 + (V / 800) ** 2 + (W / 500) ** 2), rel=1e-6)
    assert ks == pytest.approx(dt * K * math.sqrt(1 / 1000.0 ** 4 + 1 / 800.0 ** 4), rel=1e-6)
    # one cell with u = 3 U on both its faces: its Courant number wins
    c = e.field_offset("u") + (5 + 3) + e.sJ * ((4 + 3) + e.sK * 6)
    view[c] = view[c + 1] = np.float32(3 * U).view(np.uint32)
    view[so: so + 2] = 0
    ol.Program([names[i]], A[i:i + 1]).run(view)
    cfl2 = math.sqrt(view[so: so + 1].view(np.float32)[0])
    assert cfl2 == pytest.approx(dt * math.sqrt((3 * U / 1000) ** 2 + (V / 800) ** 2 + (W / 500) ** 2), rel=1e-6)


def test_set_time_step_reencodes_the_step():
    e = te_host.Run(te_host.config(**ADAPT), encode_only=True)
    names0, A0 = e.program()
    img0 = e.read(0, e.mem_words).copy()
    assert names0.count("acoustic_uv") == 2 + 2 + 4  # nsound 4 before the clock starts
    e.set_time_step(4.5, 8)
    names1, A1 = e.program()
    assert names1.count("acoustic_uv") == 3 + 4 + 8
    _, A, i = stability_op(e)
    p = e.read(int(A[i, 4 + 1]), 7)
    dtdx = f32(f32(0.5) * f32(4.5)) * f32(1 / f32(1000.0))
    assert p[:5].view(np.float32)[0] == dtdx and p[6] == 1
    e.set_time_step(6.0, 4)
    names2, A2 = e.program()
    assert names2 == names0 and np.array_equal(A2, A0)
    assert np.array_equal(e.read(0, e.mem_words), img0)
    with pytest.raises(te_host.TeError):
        e.set_time_step(4.0, 5)  # odd nsound


def test_adaptive_needs_tke():
    with pytest.raises(te_host.TeError, match="tke"):
        te_host.Run(te_host.config(**dict(ADAPT, tke=0)), encode_only=True)
