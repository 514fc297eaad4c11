"""HIP path parity at the benchmark configurations' own sizes (MI355X).

* the 192x192x80 supercell-hour workload of bench.py (TKE, Rayleigh, WENO5
  scalars, fused launches on launch shapes the autotuner chose), word for
  word against the oracle after 1 and 3 large steps;
* a 64x64x80 TKE case: at 80 levels the fused turbulence sweeps run their
  4-wave form with 3 * 80 * 64 * 4 = 61,440 bytes of LDS per block;
* the tornado-resolving grid's column (120 levels tanh-stretched from 20 m to
  300 m, dt = 0.5 s, nsound = 8 -> 15 acoustic substeps, TKE, 3 K bubble, an
  enhanced low-level hodograph) on a 64x64 plane at dx = 100 m, fused and
  unfused;
* an uneven ensemble with a 192x160x80 member, each member against the same
  configuration stepped alone, the full-size member against the oracle;
* the adjoint's reverse sweep at 128x128x60 (the adjoint workload's grid)
  against the host build's sweep over 2 steps.
"""
import sys

import numpy as np
import pytest

import adjoint_lib as al
import bench
import runs
import te_host

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare(run, ref, fused):
    """Words of the GPU run against the oracle's: profiles and every arena
    field (without the fused launches' dead scratch fields when fused)."""
    end = runs.arena_end(run) if fused else run.mem_words
    got, want = run.read(0, end), ref.M[:end].copy()
    if fused:
        keep = runs.state_mask(run)
        got[~keep] = 0
        want[~keep] = 0
    assert np.isfinite(ref.interior("th")).all() and np.isfinite(ref.interior("w")).all()
    assert np.array_equal(got, want), runs.first_difference(run, got, want)


def test_bench_workload_bit_exact(gpu):
    cfg = te_host.config(**bench.WORKLOAD)
    say("bench workload", bench.WORKLOAD)
    run = te_host.Run(cfg)
    run.init_state()
    tuned = run.autotune(reps=2, force=True)
    say("dispatches on tuned shapes", tuned)
    assert tuned > 0
    ref = runs.OracleRun(cfg)
    assert np.array_equal(run.read(0, runs.arena_end(run)), ref.M0[:runs.arena_end(run)])
    run.step(1)
    run.sync()
    ref.step(1)
    compare(run, ref, True)
    say("step 1 bit-exact")
    run.step(2)
    run.sync()
    ref.step(2)
    compare(run, ref, True)
    say("step 3 bit-exact")
    run.close()


@pytest.mark.parametrize("fuse", [1, 0])
def test_tke_80_levels(gpu, fuse):
    cfg = te_host.config(nx=64, ny=64, nz=80, dz=250.0, tke=1, initial_tke=3, rayleigh=1, fuse=fuse)
    say("64x64x80 TKE, fuse", fuse)
    run = te_host.Run(cfg)
    run.init_state()
    ref = runs.OracleRun(cfg)
    for n in (1, 2):
        run.step(n)
        run.sync()
        ref.step(n)
        compare(run, ref, fuse)
    run.close()


# the tornado-resolving grid's column: 120 tanh-stretched levels, dx = 100 m,
# dt = 0.5 s, nsound = 8 (15 acoustic substeps), TKE, +3 K bubble, and a
# quarter-circle hodograph of radius 13.8 m/s over 0-1 km (0-1 km
# storm-relative helicity R^2 pi / 2 = 299 m^2/s^2 for a storm moving with
# the circle's centre), then straight to 30 m/s at 6 km
TORNADO_COLUMN = dict(nz=120, zf=te_host.tanh_faces(), dz=0.0, dx=100.0, dy=100.0, dt=0.5, nsound=8, tke=1,
                      initial_tke=3, rayleigh=1, rayleigh_zd=18000.0, bubble_amp=3.0, bubble_rh=2000.0,
                      bubble_rv=1000.0, bubble_zc=1000.0, hodo_radius=13.8, hodo_depth=1000.0)


@pytest.mark.parametrize("fuse", [1, 0])
def test_tornado_column_stretched(gpu, fuse):
    cfg = te_host.config(nx=64, ny=64, fuse=fuse, **TORNADO_COLUMN)
    say("64x64x120 tanh-stretched column, fuse", fuse)
    run = te_host.Run(cfg)
    run.init_state()
    ref = runs.OracleRun(cfg)
    assert np.array_equal(run.read(0, runs.arena_end(run)), ref.M0[:runs.arena_end(run)])
    for n in (1, 2):
        run.step(n)
        run.sync()
        ref.step(n)
        compare(run, ref, fuse)
    run.close()


def test_ensemble_with_a_full_size_member(gpu):
    cfgs = bench.ensemble_configs(n=4)
    cfgs[0] = te_host.config(nx=192, ny=160, nz=80, dz=250.0, dt=6.0, nsound=6, tke=1, initial_tke=3, rayleigh=1,
                             sounding_seed=7, th_pert_sigma=0.5, qv_pert_sigma=0.05, shear_umax=22.0,
                             bubble_amp=2.5, seed=99)
    say("ensemble grids", [(c.nx, c.ny, c.nz) for c in cfgs])
    ens = te_host.Ensemble(cfgs)
    ens.step(2)
    ens.sync()
    for m, cfg in enumerate(cfgs):
        single = te_host.Run(cfg)
        single.init_state()
        single.step(2)
        single.sync()
        end = runs.arena_end(single)
        a, b = ens.member_words(m, 0, end), single.read(0, end)
        assert np.array_equal(a, b), (m, runs.first_difference(single, a, b))
        if m == 0:
            ref = runs.OracleRun(cfg)
            ref.step(2)
            compare(single, ref, True)
        single.close()
    ens.close()


def test_adjoint_reverse_sweep_128x128x60(gpu):
    """te_bwd on the adjoint workload's grid against the host build's reverse
    sweep (forward by the oracle, adjoint bodies on the CPU) over 2 steps."""
    steps = 2
    cfg = te_host.config(**bench.ADJOINT)
    run = te_host.Run(cfg)
    run.init_state()
    adj = te_host.Adjoint(run, every=1)
    R = runs.OracleRun(te_host.config(fuse=0, **bench.ADJOINT))
    e = R.enc
    a0, aend = e.field_offset(e.fields[0]), runs.arena_end(e)
    adj.forward(steps)
    rng = np.random.default_rng(41)
    ybar = np.zeros(aend, np.float32)
    ybar[a0:] = rng.standard_normal(aend - a0)
    for f in runs.FUSED_SCRATCH:
        if f in e.fields:
            o = e.field_offset(f)
            ybar[o: o + e.field_words] = 0
    adj.upload(0, ybar)
    adj.backward()
    dev = adj.read(0, aend).astype(np.float64)
    say("device sweep done; host sweep", steps, "steps")
    Mb = np.zeros(R._buf.size, np.float32)[runs.FRONT:]
    Mb[:aend] = ybar
    al.reverse_sweep(R.names, R.args, R.M, Mb, e.field_words, steps=steps)
    host = Mb[:aend].astype(np.float64)
    err = np.linalg.norm(dev - host) / np.linalg.norm(host)
    say("relative difference", err)
    adj.close()
    run.close()
    assert err < 1e-5
