"""Throughput of the CM1 large step on one MI355X (Mpt-steps/s, FP32).

Workload: the 192x192x80 supercell-hour configuration (dx = dy = 1 km,
dz = 250 m, dt = 6 s, nsound = 6 -> 11 acoustic substeps per large step,
1.5-order TKE, Rayleigh damping above 15 km, periodic x/y, WENO5 scalars
th qv qc and 5th-order flux-form momentum, WK82 sounding + quarter-circle
hodograph + 2 K bubble, synthetic).  The state is resident in HBM before the
timed region; a step is one full large step (a replay of its captured HIP
graph).

Prints one JSON line.  `roofline` is for the kernel with the largest device
time per step, timed live with HIP events on the stream it runs on;
`cpu_baseline` times the oracle on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "thunderengine-web_amd"))

import numpy as np  # noqa: E402

import te_host  # noqa: E402

PEAK_HBM_GBS = 8000.0
# Algorithmic bytes of one large step per interior point for this workload:
# 4 B x [P_pre + 3 (12 + 3 N_s + 3 + T_s) + 23 S + 8] with TKE (P_pre = 18,
# T_s = 4), S = 11 acoustic substeps and N_s = 4 advected scalars (th, qv, qc,
# tke): 4 x [18 + 3 (12 + 12 + 3 + 4) + 253 + 8] = 1488 (fused design)
B_ALG_STEP = 1488.0


def b_alg(n_scalars, substeps, tke):
    """Algorithmic bytes per interior point and large step (the formula above)."""
    pre, ts = (18, 4) if tke else (0, 0)
    return 4.0 * (pre + 3 * (12 + 3 * n_scalars + 3 + ts) + 23 * substeps + 8)

# kernels launched with one lane per (i, j) column that sweep k serially
COLUMN_KERNELS = {"acoustic_w_solve", "turb_scalar_vert_implicit", "turb_u_tendency", "turb_v_tendency",
                  "turb_w_tendency", "tke_transport"}
# parameters (elimination coefficients cA, cB) the *_lds launches keep in LDS
LDS_RESIDENT = {"acoustic_w_solve": (10, 11), "turb_scalar_vert_implicit": (11, 12), "turb_u_tendency": (12, 13),
                "turb_v_tendency": (12, 13), "turb_w_tendency": (11, 12), "tke_transport": (9, 10)}

WORKLOAD = dict(nx=192, ny=192, nz=80, dx=1000.0, dy=1000.0, dz=250.0, dt=6.0, nsound=6, moist=1, species=3,
                weno_scalars=1, weno_momentum=0, tke=1, rayleigh=1, rayleigh_zd=15000.0)


def kernel_io():
    """kernel -> [(param index, written)] of its buffer parameters."""
    p = os.path.join(ROOT, "thunderengine-web_amd", "kernel_io.json")
    return json.load(open(p))


def _flag(run, A, op, idx):
    """Integer word that parameter `idx` of dispatch `op` points at."""
    return int(run.read(int(A[op, 4 + idx]), 1).view(np.int32)[0])


def adv_fused_io(run, A, op):
    """(param index, read, written) of one fused advection launch (k_adv_fused
    in csrc/k_adv.hip): the advected field, the three mass fluxes, divergence,
    base-state density and map factors, then either the tendency or the fields
    the chained follow-ups (thterm, SGS adds, RK update) read and write."""
    chain, acc = _flag(run, A, op, 32), _flag(run, A, op, 12)
    io = [(i, True, False) for i in (2, 4, 6, 7, 8, 9, 10, 13, 14, 15, 16, 30)]
    if chain & 1:
        io += [(33, True, False), (34, True, False)]
    if chain & 2:
        io.append((35, True, False))
    if chain & 4:
        io.append((37, True, False))
    if chain & 64:
        io.append((42, True, False))
    if chain & (8 | 32):
        io += [(39, True, False), (40, False, True)]
    else:
        io.append((11, bool(acc) and not chain & 16, True))
    return io


def alg_bytes_per_launch(run, names, A, op):
    """Distinct arena fields a dispatch reads plus those it writes, over its
    launch grid, x 4 B (each field touched once, as a fused kernel would)."""
    name = names[op][:-4] if names[op].endswith("_lds") else names[op]
    a0 = run.field_offset(run.fields[0])
    end = a0 + run.field_words * len(run.fields)
    pts = int(A[op, 0]) * int(A[op, 1]) * int(A[op, 2])
    if name in COLUMN_KERNELS:  # one lane per column sweeps all levels
        pts *= run.nz

    def count(io):
        rd, wr = set(), set()
        for idx, read, written in io:
            off = int(A[op, 4 + idx])
            if a0 <= off < end:
                if read:
                    rd.add(off)
                if written:
                    wr.add(off)
        return len(rd) + len(wr)

    if name == "ac_pp_accumulate":  # pi' update on the interior, then the time average on the whole grid
        inner = run.nx * run.ny * run.nz
        save_x = _flag(run, A, op, 10) != 0
        pp = count([(3, 1, 0), (4, 1, 0), (5, 1, 0), (6, 1, 0), (7, 1, 1), (8, 0, 1)] + ([(9, 0, 1)] if save_x else []))
        avg = count([(2, 1, 0), (11, 1, 0), (12, 1, 0), (13, 1, 1), (14, 1, 1), (15, 1, 1)])
        return 4.0 * (pp * inner + avg * pts)
    if name == "ac_column":  # one acoustic substep after acoustic_uv, over the interior points
        inner = run.nx * run.ny * run.nz
        first = _flag(run, A, op, 33) != 0
        io = [(2, 1, 0), (3, 1, 0), (4, 1, 1), (5, 1, 0), (6, 1, 0), (7, 1, 0), (8, 1, 0), (9, 1, 0), (10, 1, 1),
              (11, 0, 1), (13, not first, 1), (14, not first, 1), (15, 1, 1)]
        if _flag(run, A, op, 16) != 0:
            io.append((12, 0, 1))
        return 4.0 * count(io) * inner
    if name == "bc_periodic_xy":  # each halo cell reads its interior image and writes itself
        n = _flag(run, A, op, 9)
        io = [(1 + q, True, True) for q in range(n)]
    elif name == "zero_fields":
        io = [(1 + q, False, True) for q in range(_flag(run, A, op, 9))]
    elif name == "add_fields":
        n = _flag(run, A, op, 13)
        io = [(1 + q, True, False) for q in range(n)] + [(5 + q, True, True) for q in range(n)]
    elif name == "copy_fields":
        n = _flag(run, A, op, 9)
        io = [(1 + q, True, False) for q in range(n)] + [(5 + q, False, True) for q in range(n)]
    elif name == "tendency_reset_add":
        io = [(1 + q, False, True) for q in range(3)] + [(4 + q, True, False) for q in range(3)]
    elif name == "adv_fused":
        io = adv_fused_io(run, A, op)
    elif name == "massflux_div":
        io = [(1, True, False), (2, True, False), (3, True, False), (6, False, True), (7, False, True),
              (8, False, True), (10, False, True)]
    elif name.endswith("_xy") and name[:-3] in kernel_io():  # an update, then its outputs' periodic halo cells
        halo = ((run.nx + 7) * (run.ny + 7) - run.nx * run.ny) * (run.nz + 1)
        return 4.0 * (count(kernel_io()[name[:-3]]) * pts + halo * (4 if name == "turb_km_kh_xy" else 1))
    else:
        io = kernel_io().get(name)
    if io is None:
        return None
    if names[op].endswith("_lds"):
        io = [r for r in io if r[0] not in LDS_RESIDENT.get(name, ())]
    return 4.0 * count(io) * pts


def profiles_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (profiles/hbm_traffic_192x192x80.txt: FETCH_SIZE + WRITE_SIZE per dispatch,
    each scaled by the copy-kernel calibration), or None."""
    p = os.path.join(ROOT, "profiles", "hbm_traffic_192x192x80.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        f = line.split()
        if len(f) >= 5 and f[0] == kernel:
            return float(f[4])
    return None


def calibration_copy(words=1 << 26):
    """One plain copy of `words` 32-bit words between two fresh buffers (the
    PMC runs read FETCH_SIZE / WRITE_SIZE of this dispatch: 4*words bytes each)."""
    A = np.zeros((1, 64), np.uint32)
    A[0, 0], A[0, 4], A[0, 5] = words, 0, words
    r = te_host.Replayer(2 * words, ["_copy"], A)
    r.run(1)
    r.sync()
    r.close()


def ensemble_configs(n=32, seed=20260102):
    """Seeded uneven ensemble: nx, ny in {96, 128, 160, 192}, nz in {40, 60, 80}
    with dz = 20 km / nz; theta' N(0, 0.5 K) smooth in z, qv x (1 + N(0, 0.05)),
    0-6 km shear U(10, 40) m/s, bubble U(1, 3) K; TKE as the main workload."""
    rng = np.random.default_rng(seed)
    cfgs = []
    for m in range(n):
        nx, ny = (int(v) for v in rng.choice([96, 128, 160, 192], 2))
        nz = int(rng.choice([40, 60, 80]))
        cfgs.append(te_host.config(nx=nx, ny=ny, nz=nz, dz=20000.0 / nz, dt=6.0, nsound=6, tke=1, rayleigh=1,
                                   rayleigh_zd=15000.0, sounding_seed=seed + m, th_pert_sigma=0.5,
                                   qv_pert_sigma=0.05, shear_umax=float(rng.uniform(10, 40)),
                                   bubble_amp=float(rng.uniform(1, 3)), seed=seed + 1000 + m))
    return cfgs


def ensemble_bench(steps, warmup, tune=True):
    cfgs = ensemble_configs()
    ens = te_host.Ensemble(cfgs)
    tuned = ens.autotune() if tune else 0  # per member, untimed
    ens.step(warmup)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    wall = time.perf_counter() - t0
    out = {"members": len(cfgs), "points": ens.points, "steps": steps, "autotuned_ops": tuned,
           "value": round(ens.points * steps / wall / 1e6, 2), "unit": "Mpt-steps/s",
           "grids": sorted({"%dx%dx%d" % (c.nx, c.ny, c.nz) for c in cfgs})}
    ens.close()
    return out


# Adjoint workload: 128 x 128 x 60, dx = dy = 1 km, dz = 333 m, dt = 6 s,
# nsound 6, WENO scalars, theta qv qc, Rayleigh damping, no turbulence;
# 60 large steps with a checkpoint every 10.
ADJOINT = dict(nx=128, ny=128, nz=60, dz=20000.0 / 60, dt=6.0, nsound=6)


def adjoint_bench(steps=60, every=10):
    """Cost of the gradient of J = sum of w^2 over 0-8 km in the 20 x 20 km
    column block around the bubble after `steps` steps, with respect to the
    initial theta', u, v and qv: the checkpointed forward plus the reverse
    sweep, against the plain forward of the same steps (graph replays)."""
    cfg = te_host.config(**ADJOINT)
    run = te_host.Run(cfg)
    run.init_state()
    run.step(2)
    run.sync()
    run.init_state()
    t0 = time.perf_counter()
    run.step(steps)
    run.sync()
    plain = time.perf_counter() - t0
    run.init_state()
    adj = te_host.Adjoint(run, every=every)
    t0 = time.perf_counter()
    adj.forward(steps)
    fwd = time.perf_counter() - t0
    nwords = run.field_offset(run.fields[0]) + run.field_words * len(run.fields)  # profiles + arena
    c = cfg.nx // 2
    top = int(8000.0 / cfg.dz)
    box = te_host.Box(run, i=(c - 10, c + 10), j=(c - 10, c + 10), k=(1, top + 1))
    state = run.read(0, nwords).view(np.float32)
    seed = np.zeros(nwords, np.float32)
    seed[box.words] = 2.0 * state[box.words]
    adj.upload(0, seed)
    t0 = time.perf_counter()
    adj.backward()
    bwd = time.perf_counter() - t0
    ctl = te_host.Controls(run, ("th", "u", "v", "qv"))
    g = ctl.pull(adj.read(0, nwords)).reshape(4, -1)
    fwd_ms, bwd_ms = adj.last_ms()
    adj.close()
    run.close()
    return {"workload": "128x128x60 dz=333m dt=6s nsound=6, WENO scalars, no turbulence, %d steps, "
                        "checkpoint every %d" % (steps, every),
            "plain_fwd_s": round(plain, 4), "fwd_ckpt_s": round(fwd, 4), "bwd_s": round(bwd, 4),
            "cost_vs_plain_fwd": round((fwd + bwd) / plain, 2),
            "J": float(np.sum(state[box.words].astype(np.float64) ** 2)),
            "grad_norms": {k: float(np.linalg.norm(g[q])) for q, k in enumerate(("theta'", "u", "v", "qv"))},
            "device_ms": {"fwd_ckpt": round(fwd_ms, 2), "bwd": round(bwd_ms, 2)}}


# The tornado-resolving grid's forward step: 512 x 512 x 120 at dx = dy = 100 m,
# tanh-stretched dz from 20 m (297 m at 16 km), dt = 0.5 s, nsound = 8 (15
# acoustic substeps), TKE, +3 K bubble, quarter-circle hodograph of radius
# 13.8 m/s over 0-1 km (0-1 km storm-relative helicity ~ 300 m^2/s^2).
TORNADO = dict(nx=512, ny=512, nz=120, zf=te_host.tanh_faces(), dz=0.0, dx=100.0, dy=100.0, dt=0.5, nsound=8,
               tke=1, initial_tke=3, rayleigh=1, rayleigh_zd=18000.0, bubble_amp=3.0, bubble_rh=2000.0,
               bubble_rv=1000.0, bubble_zc=1000.0, hodo_radius=13.8, hodo_depth=1000.0)


def tornado_bench(steps, warmup, tune=True):
    """Forward Mpt-steps/s of the tornado-resolving grid (graph replays, state resident)."""
    cfg = te_host.config(**TORNADO)
    run = te_host.Run(cfg)
    run.init_state()
    if tune:
        run.autotune(reps=2)
    run.step(warmup)
    run.sync()
    t0 = time.perf_counter()
    run.step(steps)
    run.sync()
    wall = time.perf_counter() - t0
    pts = cfg.nx * cfg.ny * cfg.nz
    value = pts * steps / wall / 1e6
    balg = b_alg(4, 15, True)
    out = {"workload": "512x512x120 tanh dz 20-300 m, dx 100 m, dt 0.5 s, nsound 8 (15 substeps), TKE, 3 K bubble",
           "steps": steps, "value": round(value, 2), "unit": "Mpt-steps/s", "ms_per_step": round(wall * 1e3 / steps, 3),
           "alg_bytes_per_pt_step": balg, "hbm_frac": round(value * 1e6 * balg / 8.0e12, 4),
           "state_GB": round(run.mem_words * 4 / 1e9, 2)}
    run.close()
    return out


def cpu_baseline(cfg, seconds=15.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    import runs
    ref = runs.OracleRun(cfg)
    ref.step(1)  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        ref.step(1)
        n += 1
        dt = time.perf_counter() - t0
        if dt > seconds or n >= 20:
            break
    pts = cfg.nx * cfg.ny * cfg.nz
    return {"value": round(pts * n / dt / 1e6, 3), "unit": "Mpt-steps/s", "cores": int(oracle_lib.lib().oracle_threads()),
            "kind": "port",
            "sample": "%d large steps of the same 192x192x80 workload after 1 warm-up step, oracle (C++ -O2, OpenMP)" % n,
            "reference_cpu_paths": {"WGSL step on a CPU WebGPU adapter": "unavailable (no WebGPU implementation on the "
                                    "host, and the reference's ES2020 modules do not load in its Node 12)",
                                    "CM1 r20.2 Fortran FP32": "unavailable (no Fortran compiler, CM1 source not "
                                    "in the reference)"}}


def log(*a):
    print("bench:", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--calib-copy", action="store_true", help="also launch the PMC calibration copy")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 32-member uneven ensemble (members stepped together)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint cost (128x128x60, 60 steps, checkpoint every 10)")
    ap.add_argument("--no-tornado", action="store_true",
                    help="skip the 512x512x120 tornado-resolving grid's forward step")
    ap.add_argument("--ensemble-steps", type=int, default=120, help="large steps of the ensemble timing")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (no te_autotune before the timed steps)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy:
        calibration_copy()
    cfg = te_host.config(**WORKLOAD)
    run = te_host.Run(cfg)
    run.init_state()
    pts = cfg.nx * cfg.ny * cfg.nz
    tuned, tune_s, tune_ms = 0, 0.0, (0.0, 0.0)
    if not args.no_autotune:  # untimed: shapes chosen on the t=0 state, memory restored
        t0 = time.perf_counter()
        tuned = run.autotune(reps=5)
        tune_s = time.perf_counter() - t0
        tune_ms = run.autotune_ms()
    if args.warmup > 0:
        run.step(args.warmup)
    run.sync()
    t0 = time.perf_counter()
    run.step(args.steps)
    run.sync()
    wall = time.perf_counter() - t0
    dev_ms = run.last_run_ms()
    ms_per_step = wall * 1e3 / args.steps
    value = pts * args.steps / wall / 1e6
    log("%d timed steps: %.3f ms/step, %.1f Mpt-steps/s" % (args.steps, ms_per_step, value))

    # dominant kernel: largest device time per step, per-op HIP events
    names, A = run.program()
    op_ms = run.profile_ops(args.profile_steps) / args.profile_steps
    per_kernel = {}
    for i, n in enumerate(names):
        t, c = per_kernel.get(n, (0.0, 0))
        per_kernel[n] = (t + float(op_ms[i]), c + 1)
    dom = max(per_kernel, key=lambda k: per_kernel[k][0])
    dom_ops = [i for i, n in enumerate(names) if n == dom]
    avg_ms = per_kernel[dom][0] / len(dom_ops)
    alg_all = [alg_bytes_per_launch(run, names, A, i) for i in range(len(names))]
    alg = [alg_all[i] for i in dom_ops]
    alg_b = float(np.mean(alg)) if all(x is not None for x in alg) else None
    kernels_alg_pt = sum(x for x in alg_all if x) / pts
    achieved = alg_b / (avg_ms * 1e-3) / 1e9 if alg_b else None
    traffic = profiles_traffic(dom)
    roof = {"bound": "hbm", "kernel": dom, "launches_per_step": len(dom_ops), "avg_launch_ms": round(avg_ms, 5),
            "alg_bytes_per_launch": alg_b,
            "achieved": round(achieved, 1) if achieved else None, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4) if achieved else None, "traffic": traffic}
    step_gbs = value * 1e6 * B_ALG_STEP / 1e9
    top_frac = {}
    for k, (t, _) in sorted(per_kernel.items(), key=lambda kv: -kv[1][0])[:8]:
        b = [alg_all[i] for i, n in enumerate(names) if n == k]
        if t > 0 and all(x is not None for x in b):
            top_frac[k] = round(sum(b) / (t * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)

    out = {
        "metric": "Mpt-steps/s FP32 at 192^2x80 (%HBM peak)",
        "value": round(value, 2), "unit": "Mpt-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "device_ms_per_step": round(dev_ms / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "supercell hour 192x192x80 dz=250m dt=6s nsound=6 TKE (synthetic WK82 + bubble)",
                   "nx": cfg.nx, "ny": cfg.ny, "nz": cfg.nz, "substeps_per_step": 11, "ops_per_step": len(names),
                   "parallelism": "single GPU", "autotuned_ops": tuned, "autotune_s": round(tune_s, 2),
                   "autotune_eager_ms": [round(tune_ms[0], 4), round(tune_ms[1], 4)]},
        "step_hbm": {"alg_bytes_per_pt_step": B_ALG_STEP, "achieved_GBs": round(step_gbs, 1),
                     "frac": round(step_gbs / PEAK_HBM_GBS, 4),
                     "launch_alg_bytes_per_pt_step": round(kernels_alg_pt, 1)},
        "roofline": roof,
        "top_kernels_ms_per_step": {k: round(v[0], 4) for k, v in
                                    sorted(per_kernel.items(), key=lambda kv: -kv[1][0])[:8]},
        "top_kernels_hbm_frac": top_frac,
    }
    run.close()
    log("per-launch profile done, dominant kernel", dom)
    if not args.no_ensemble:
        ens = ensemble_bench(args.ensemble_steps, 2, tune=not args.no_autotune)
        ens["per_point_vs_plain"] = round(ens["value"] / value, 3)
        out["ensemble"] = ens
        log("ensemble", ens["value"], "Mpt-steps/s")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
        log("adjoint done")
    if not args.no_tornado:
        out["tornado_grid"] = tornado_bench(10, 2, tune=not args.no_autotune)
        log("tornado grid", out["tornado_grid"]["value"], "Mpt-steps/s")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg)
        log("cpu baseline", out["cpu_baseline"]["value"], "Mpt-steps/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
