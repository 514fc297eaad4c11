"""Whole-step invariants of the oracle on periodic configurations (CPU).

The recorded tape is an open-boundary case, so the periodic halo kernels and
the periodic index arithmetic of the large step are pinned here instead:

* halo cells hold exact periodic images of the interior after a step;
* the step commutes with a circular shift of the domain: stepping a shifted
  initial state gives, bit for bit, the shifted result (nx a multiple of 4,
  so the FMA-order switch of acoustic_prepare is the same for every column).
  With TKE this holds for one step only: the reference never refreshes the
  halo of rho after update_rho_prs (web/core/timestep.js:2132-2141 writes the
  interior; turb_rho_faces and turb_def_centers, timestep.js:1074-1090, read
  rho at i, j = -1 and -2), so from step 2 the turbulence at the domain edge
  sees t=0 halo values, which a shift does not commute with.
"""
import numpy as np
import pytest

import runs
import te_host

PROGNOSTIC = ["th", "u", "v", "w", "pp3d", "qv", "tke"]


def interior_wrap(f, nx, ny):
    inner = f[:, 3:3 + ny, 3:3 + nx]
    return np.pad(inner, ((0, 0), (3, 4), (3, 4)), mode="wrap")


@pytest.mark.parametrize("kw", [dict(), dict(tke=1, initial_tke=3, rayleigh=1)])
def test_periodic_halos_after_step(kw):
    r = runs.OracleRun(te_host.config(nx=24, ny=20, nz=12, **kw))
    r.step(2)
    e = r.enc
    for name in ("th", "u", "v", "w", "qv", "pp3d"):
        f = r.field(name)[: e.nz]
        want = interior_wrap(f, e.nx, e.ny)
        assert np.array_equal(f.view(np.uint32), want.view(np.uint32)), name


@pytest.mark.parametrize("kw,steps", [(dict(), 3), (dict(weno_momentum=1, nsound=4), 2),
                                      (dict(tke=1, initial_tke=3, weno_momentum=1, rayleigh=1), 1)])
def test_step_commutes_with_circular_shift(kw, steps):
    cfg = te_host.config(nx=24, ny=20, nz=12, **kw)
    a = runs.OracleRun(cfg)
    b = runs.OracleRun(cfg)
    e = a.enc
    sx, sy = 7, 5
    a0 = e.field_offset(e.fields[0])
    n = e.sJ * e.sK * (e.nz + 1)
    for q in range(len(e.fields)):
        o = a0 + q * e.field_words
        f = b.M[o: o + n].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)
        inner = np.roll(f[:, 3:3 + e.ny, 3:3 + e.nx], (sy, sx), axis=(1, 2))
        f[:] = np.pad(inner, ((0, 0), (3, 4), (3, 4)), mode="wrap")
    a.step(steps)
    b.step(steps)
    for name in PROGNOSTIC:
        if name not in e.fields:
            continue
        fa = np.roll(a.field(name)[:, 3:3 + e.ny, 3:3 + e.nx], (sy, sx), axis=(1, 2))
        fb = b.field(name)[:, 3:3 + e.ny, 3:3 + e.nx]
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32)), name


@pytest.mark.parametrize("kw", [dict(), dict(tke=1, initial_tke=3, rayleigh=1), dict(moist=0, nsound=4)])
def test_fused_program_equals_reference_program(kw):
    """The library's fused dispatch list (merged periodic halos, reordered
    halo fills, the one-launch acoustic substep) leaves the same bits as the
    reference's list in the profiles and every field the step carries (the
    fused launches' scratch fields may differ)."""
    cfg = te_host.config(nx=24, ny=20, nz=12, **kw)
    a = runs.OracleRun(cfg)
    b = runs.OracleRun(cfg, fused=True)
    assert len(b.names) < len(a.names)
    a.step(2)
    b.step(2)
    end = runs.arena_end(a.enc)
    assert runs.arena_end(b.enc) == end  # parameter bytes follow the arena and may differ
    keep = runs.state_mask(a.enc)
    wa, wb = a.M[:end].copy(), b.M[:end].copy()
    wa[~keep] = 0
    wb[~keep] = 0
    assert np.array_equal(wa, wb), runs.first_difference(a, wa, wb)


@pytest.mark.parametrize("kw", [dict(tke=1, initial_tke=3, rayleigh=1), dict(open_x=1, open_y=1, tke=1, initial_tke=3)])
def test_fused_scratch_fields_are_dead(kw):
    """Garbage in any FUSED_SCRATCH field at the start of a step changes no
    other field at its end: fused launches may leave them different."""
    cfg = te_host.config(nx=20, ny=16, nz=10, **kw)
    base = runs.OracleRun(cfg)
    base.step(1)
    e = base.enc
    n = e.sJ * e.sK * (e.nz + 1)
    junk = np.frombuffer(np.full(n, 123.456, np.float32).tobytes(), np.uint32)
    for f in runs.FUSED_SCRATCH:
        if f not in e.fields:
            continue
        r = runs.OracleRun(cfg)
        o = e.field_offset(f)
        r.M[o: o + n] = junk
        r.step(1)
        for g in e.fields:
            if g == f:
                continue
            og = e.field_offset(g)
            assert np.array_equal(r.M[og: og + n], base.M[og: og + n]), (f, g)


def test_fp64_build_is_the_same_step():
    """The FP64 instantiation of the oracle follows the FP32 one to FP32
    rounding over a few steps (a kernel whose float and double forms
    differed would show as an O(1) difference)."""
    import fp64_error
    res = fp64_error.run(te_host.config(nx=24, ny=20, nz=12, tke=1, initial_tke=3, rayleigh=1), [1, 3])
    for steps, row in res.items():
        for label, e in row.items():
            assert e["max_abs"] <= 1e-3 * max(1.0, e["max_abs_fp64"]), (steps, label, e)


def test_perturbed_sounding_is_seeded_and_smooth():
    """Sounding perturbations (ensemble members) are reproducible from the
    seed, leave the default sounding untouched, and theta' is smooth in z."""
    base = runs.OracleRun(te_host.config(nx=16, ny=16, nz=40, dz=500.0))
    a = runs.OracleRun(te_host.config(nx=16, ny=16, nz=40, dz=500.0, sounding_seed=5, th_pert_sigma=0.5,
                                      qv_pert_sigma=0.05))
    b = runs.OracleRun(te_host.config(nx=16, ny=16, nz=40, dz=500.0, sounding_seed=5, th_pert_sigma=0.5,
                                      qv_pert_sigma=0.05))
    c = runs.OracleRun(te_host.config(nx=16, ny=16, nz=40, dz=500.0, sounding_seed=6, th_pert_sigma=0.5))
    e = base.enc
    th0 = lambda r: r.M[e.region_offset("th0"): e.region_offset("th0") + e.nz].view(np.float32).astype(np.float64)
    assert np.array_equal(a.M0, b.M0)
    d = th0(a) - th0(base)
    assert 0.1 < d.std() < 1.5 and np.abs(d).max() < 3.0
    assert np.abs(np.diff(d)).max() < 0.5 * np.abs(d).max() + 1e-9   # smooth over ~1 km (2 levels)
    assert not np.array_equal(th0(a), th0(c))
    default = runs.OracleRun(te_host.config(nx=16, ny=16, nz=40, dz=500.0, th_pert_sigma=0.0))
    assert np.array_equal(default.M0, base.M0)
