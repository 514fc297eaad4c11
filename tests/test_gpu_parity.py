"""HIP path parity (MI355X): bit-exact against the oracle and the reference's
recorded hashes, through the C ABI (libte.so).

* the recorded preset tapes (supercell, squall line, stretched grid) replayed on the GPU, dispatch by dispatch,
  against the region hashes the reference recorded on its own GPU;
* programs of the encoder on synthetic configurations (the 64x64x40 case,
  TKE, open boundaries, dry, WENO momentum, explicit vertical diffusion) run
  for one and several large steps on the GPU (several steps replay a captured
  HIP graph) and compared word for word with the oracle on the same memory.
"""
import os
import sys

import numpy as np
import pytest

import golden_check
import runs
import tapeplan
import te_host

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("preset", tapeplan.PRESETS)
def test_preset_tape_region_hashes(gpu, preset):
    P = golden_check.plan(preset)
    have = set(te_host.kernel_names())
    names, A = golden_check.program_for(P, have)
    say(preset, "tape on GPU:", len(names), "dispatches, memory words", P.mem_words)
    r = te_host.Replayer(P.mem_words, names, A)
    r.upload(0, P.memory())
    ok, bad, skipped, mism = golden_check.check(P, lambda i: r.run(1, i, i), r.region_hash, have)
    r.close()
    say("ok", ok, "bad", bad, "skipped", skipped)
    assert bad == 0, mism[:10]
    assert ok >= {"supercell": 825, "squall_line": 815, "supercell_stretched": 52}[preset]


HURRICANE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "soundings",
                         "hurricane_3d.input_sounding")

CONFIGS = {
    "A_64x64x40": dict(),
    # the caller's sounding: the reference's hurricane_3d input_sounding (isnd = 7)
    "sounding_hurricane": dict(nx=40, ny=32, nz=24, dz=600.0, tke=1, initial_tke=3, rayleigh=1, umove=2.0,
                               input_sounding=open(HURRICANE).read(), sounding_winds=1),
    "tke_rayleigh": dict(nx=48, ny=40, nz=24, tke=1, rayleigh=1, initial_tke=3),
    "open_tke_heating": dict(nx=40, ny=48, nz=20, open_x=1, open_y=1, tke=1, initial_tke=3, umove=12.5, vmove=3.0,
                             dissipative_heating=1, rayleigh=1),
    "dry_explicit_vdiff": dict(nx=36, ny=36, nz=20, moist=0, tke=1, initial_tke=3, explicit_vdiff=1),
    "weno_momentum": dict(nx=40, ny=36, nz=24, weno_momentum=1),
    "species7_ragged": dict(nx=37, ny=29, nz=17, species=7, tke=1, initial_tke=4, nsound=4, dt=4.0),
    "open_mass_conservation": dict(nx=30, ny=26, nz=16, open_x=1, open_y=1, tke=1, initial_tke=3, umove=5.0,
                                   mass_conservation=1, rayleigh=1),
    # stretched horizontal grids: x 14 + 20 + 14 cells from 500 m to 2 km on both
    # sides, y 24 + 16 cells stretched on the north side only; substeps from the
    # 500 m finest spacing (nsound 8 -> 15 substeps)
    "stretched_tke": dict(nx=48, ny=40, nz=20, nsound=0, tke=1, initial_tke=3, rayleigh=1, stretch_x=1,
                          dx_inner=500.0, dx_outer=2000.0, nos_x_len=10000.0, tot_x_len=45000.0, stretch_y=2,
                          dy_inner=500.0, dy_outer=2000.0, nos_y_len=12000.0, tot_y_len=32000.0),
    # the supercell_stretched preset's options on a 36 x 36 grid (12 + 12 + 12
    # cells per axis, 500 m centre, 2 km edges), open sides
    "stretched_open_supercell": dict(nx=36, ny=36, nz=20, dx=1250.0, dy=1250.0, dt=3.0, nsound=0, open_x=1,
                                     open_y=1, tke=1, initial_tke=3, umove=12.5, vmove=3.0, dissipative_heating=1,
                                     rayleigh=1, mass_conservation=1, stretch_x=1, dx_inner=500.0, dx_outer=2000.0,
                                     nos_x_len=6000.0, tot_x_len=36000.0, stretch_y=1, dy_inner=500.0,
                                     dy_outer=2000.0, nos_y_len=6000.0, tot_y_len=36000.0),
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
@pytest.mark.parametrize("steps", [1, 3])
@pytest.mark.parametrize("fuse", [1, 0])
def test_config_bit_exact(gpu, name, steps, fuse):
    """fuse=0 runs the reference's dispatch list (whole memory must match);
    fuse=1 the library's fused launches (profiles and every arena field but
    the dead scratch fields must match)."""
    cfg = te_host.config(fuse=fuse, **CONFIGS[name])
    say("config", name, "steps", steps, "fuse", fuse)
    ref = runs.OracleRun(cfg)
    run = te_host.Run(cfg)
    run.init_state()
    end = runs.arena_end(run) if fuse else run.mem_words
    M0 = run.read(0, end)
    assert np.array_equal(M0, ref.M0[:end]), runs.first_difference(run, M0, ref.M0[:end])
    run.step(steps)
    run.sync()
    got = run.read(0, end)
    ref.step(steps)
    want = ref.M[:end].copy()
    assert np.isfinite(ref.interior("th")).all()
    if fuse:
        keep = runs.state_mask(run)
        got[~keep] = 0
        want[~keep] = 0
    assert np.array_equal(got, want), runs.first_difference(run, got, want)
    run.close()


@pytest.mark.parametrize("name", ["species7_ragged", "tke_rayleigh"])
@pytest.mark.parametrize("fuse", [1, 0])
def test_autotuned_launch_shapes_bit_exact(gpu, name, fuse):
    """te_autotune leaves the memory as it found it, and the program with the
    shapes it chose still matches the oracle word for word."""
    cfg = te_host.config(fuse=fuse, **CONFIGS[name])
    say("autotune", name, "fuse", fuse)
    run = te_host.Run(cfg)
    run.init_state()
    run.step(1)  # a graph exists: the tuner must capture it again
    before = run.read(0, run.mem_words)
    tuned = run.autotune(reps=2, force=True)
    say("tuned dispatches", tuned, "eager step ms default / tuned", run.autotune_ms())
    assert tuned > 0
    assert np.array_equal(run.read(0, run.mem_words), before)
    run.step(2)
    run.sync()
    ref = runs.OracleRun(cfg)
    ref.step(3)
    end = runs.arena_end(run) if fuse else run.mem_words
    got, want = run.read(0, end), ref.M[:end].copy()
    if fuse:
        keep = runs.state_mask(run)
        got[~keep] = 0
        want[~keep] = 0
    assert np.array_equal(got, want), runs.first_difference(run, got, want)
    run.close()


def test_ensemble_members_match_single_runs(gpu):
    """Members of uneven grids and perturbed soundings stepped together on
    their own streams equal the same configurations stepped alone, and one
    perturbed member equals the oracle."""
    kws = [dict(nx=32, ny=24, nz=16, dz=1250.0, tke=1, initial_tke=3, sounding_seed=11, th_pert_sigma=0.5,
                qv_pert_sigma=0.05, shear_umax=17.0, bubble_amp=1.5),
           dict(nx=40, ny=32, nz=20, dz=1000.0, sounding_seed=12, th_pert_sigma=0.5, qv_pert_sigma=0.05,
                shear_umax=35.0, bubble_amp=2.5),
           dict(nx=24, ny=40, nz=12, dz=1666.6667, tke=1, initial_tke=4, sounding_seed=13, th_pert_sigma=0.5)]
    cfgs = [te_host.config(**kw) for kw in kws]
    ens = te_host.Ensemble(cfgs)
    assert len(ens) == 3 and ens.points == sum(k["nx"] * k["ny"] * k["nz"] for k in kws)
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
            keep = runs.state_mask(single)
            want = ref.M[:end].copy()
            got = a.copy()
            got[~keep] = 0
            want[~keep] = 0
            assert np.array_equal(got, want), runs.first_difference(single, got, want)
    ens.close()


def test_graph_replays_match_eager(gpu):
    """Steps replayed from the captured graph equal eager single steps."""
    cfg = te_host.config(nx=32, ny=32, nz=16, tke=1, initial_tke=3)
    a, b = te_host.Run(cfg), te_host.Run(cfg)
    a.init_state(); b.init_state()
    for rep in range(3):
        a.step(2)            # graph
        b.step(1); b.step(1)  # eager
        a.sync(); b.sync()
        wa, wb = a.read(0, a.mem_words), b.read(0, b.mem_words)
        assert np.array_equal(wa, wb), (rep, runs.first_difference(a, wa, wb))
