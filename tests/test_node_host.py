"""The JavaScript host (Node N-API addon + thunderengine.js).

CPU: its tape planner agrees with the reference's replay plan restated in
tests/tapeplan.py, and its buildRun encodes the same program as the C ABI.
GPU: it replays the recorded supercell tape with per-dispatch hashes equal to
the reference's, and steps configurations bit-exactly against the oracle.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import golden_check
import oracle_lib as ol
import runs
import tapeplan
import te_host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests")
SCRIPT = os.path.join(ROOT, "tests", "js", "check_host.js")
ADDON = os.path.join(ROOT, "thunderengine-web_amd", "lib", "te_napi.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the N-API addon is not available")


def node(*args, timeout=300):
    out = subprocess.run(["node", SCRIPT] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stderr[-2000:]
    return out.stdout


def fnv(words):
    h = 0x811C9DC5
    for w in np.asarray(words, np.uint32).tolist():
        h = ((h ^ w) * 0x01000193) & 0xFFFFFFFF
    return h


def test_tape_plan_matches_reference_plan():
    got = json.loads(node("plan"))
    P = golden_check.plan()
    names, A = P.args()
    assert got["memWords"] == P.mem_words
    assert got["ops"] == len(names)
    assert got["slots"] == len(P.slots) == 1010
    assert got["argsFnv"] == ol.fnv1a(A.reshape(-1))
    assert got["slotFnv"] == ol.fnv1a(np.array([s[2] for s in P.slots], np.uint32))


def test_addon_refuses_bad_offsets():
    got = json.loads(node("ranges"))
    assert got["ok"] and all(got["refused"]), got


def test_build_run_encodes_like_the_c_abi():
    conf = dict(nx=24, ny=20, nz=12, tke=1, rayleigh=1)
    got = json.loads(node("encode", json.dumps(conf)))
    r = te_host.Run(te_host.config(**conf), encode_only=True)
    names, A = r.program()
    assert got["names"] == names
    assert got["argsFnv"] == ol.fnv1a(A.reshape(-1))
    assert got["memWords"] == r.mem_words


def test_preset_mapping():
    conf = {"param0": {"nx": 24, "ny": 20, "nz": 12}, "param1": {"dx": 1000.0, "dy": 1000.0, "dz": 500.0, "dtl": 6.0},
            "param2": {"advwenos": 2, "advwenov": 0, "sgsmodel": 1, "doimpl": 1, "irdamp": 1, "wbc": 1, "ebc": 1,
                       "sbc": 1, "nbc": 1, "imoist": 1, "ptype": 0, "idiss": 1, "imove": 0},
            "param3": {"kdiv": 0.1, "alph": 0.6, "zd": 15000.0, "rdalpha": 0.0033333333333}}
    got = json.loads(node("encode", json.dumps(conf)))
    r = te_host.Run(te_host.config(nx=24, ny=20, nz=12, tke=1, rayleigh=1, dissipative_heating=1), encode_only=True)
    names, A = r.program()
    assert got["names"] == names
    assert got["argsFnv"] == ol.fnv1a(A.reshape(-1))


def test_stretched_preset_mapping():
    """stretch_x / stretch_y and their lengths reach te_config from a namelist
    preset; the encoded memory (hmetric: map factors, faces) is the C ABI's."""
    conf = {"param0": {"nx": 48, "ny": 40, "nz": 12}, "param1": {"dx": 1000.0, "dy": 800.0, "dz": 500.0, "dtl": 6.0},
            "param2": {"wbc": 2, "ebc": 2, "sbc": 2, "nbc": 2},
            "param4": {"stretch_x": 1, "dx_inner": 500.0, "dx_outer": 2000.0, "nos_x_len": 10000.0,
                       "tot_x_len": 45000.0},
            "param5": {"stretch_y": 2, "dy_inner": 500.0, "dy_outer": 2000.0, "nos_y_len": 12000.0,
                       "tot_y_len": 32000.0}}
    got = json.loads(node("encode", json.dumps(conf)))
    r = te_host.Run(te_host.config(nx=48, ny=40, nz=12, dx=1000.0, dy=800.0, open_x=1, open_y=1, stretch_x=1,
                                   dx_inner=500.0, dx_outer=2000.0, nos_x_len=10000.0, tot_x_len=45000.0,
                                   stretch_y=2, dy_inner=500.0, dy_outer=2000.0, nos_y_len=12000.0,
                                   tot_y_len=32000.0), encode_only=True)
    names, A = r.program()
    assert got["names"] == names
    assert got["argsFnv"] == ol.fnv1a(A.reshape(-1))
    assert got["memFnv"] == ol.fnv1a(r.read(0, r.mem_words))
    flat = te_host.config(nx=48, ny=40, nz=12, dx=1000.0, dy=800.0, open_x=1, open_y=1)
    assert got["memFnv"] != ol.fnv1a(te_host.Run(flat, encode_only=True).read(0, r.mem_words))


def test_adaptive_clock_binding():
    """The JavaScript AdaptiveTimeStep (N-API) takes the C ABI clock's steps."""
    rng = np.random.default_rng(5)
    cfl = [float(np.float32(x)) for x in rng.uniform(0.1, 1.6, 30)]
    ks = [float(np.float32(x)) for x in rng.uniform(0.0, 0.12, 30)]
    arg = dict(dtl=6.0, dmin=750.0, frequency={"tape": 90.0, "restart": 0.0, "stat": 30.0}, cfl=cfl, ks=ks)
    got = json.loads(node("clock", json.dumps(arg)))
    c = te_host.Clock(6.0, 750.0, freq=arg["frequency"])
    c.start(cfl[0], ks[0])
    st = c.state()
    want = [[st["dt"], st["nsound"], st["mtime"], ",".join(st["fired_kinds"])]]
    for i in range(1, len(cfl)):
        c.begin_step()
        fired = ",".join(c.state()["fired_kinds"])
        c.end_step(cfl[i], ks[i])
        st = c.state()
        want.append([st["dt"], st["nsound"], st["mtime"], fired])
    assert got == want


TAPE_CONFS = {
    "tke_open": (dict(nx=20, ny=16, nz=10, tke=1, initial_tke=3, open_x=1, open_y=1, umove=5.0, rayleigh=1), None),
    "sounding": (dict(nx=16, ny=12, nz=20, dz=400.0, tke=1, initial_tke=3, rayleigh=1),
                 os.path.join(HERE, "golden", "soundings", "hurricane_3d.input_sounding")),
}


@pytest.mark.parametrize("name", sorted(TAPE_CONFS))
def test_tape_export_replays_like_the_program(name, tmp_path):
    """run.tape(2) + run.tapeState(): the replay plan of the exported tape
    (replay.js:46-121, the plan recorded tapes go through) stepped by the
    oracle gives the arena of the oracle stepping the program directly."""
    conf, snd = TAPE_CONFS[name]
    args = ["tape", json.dumps(conf), str(tmp_path)] + ([snd] if snd else [])
    info = json.loads(node(*args))
    tape = json.load(open(tmp_path / "tape.json"))
    assert info["steps"] == 2 and len(tape["steps"]) == 2
    io = json.load(open(os.path.join(ROOT, "thunderengine-web_amd", "kernel_io.json")))
    params = {k: [[i, "ptr", int(w)] for i, r, w in v] for k, v in io.items()}
    P = tapeplan.Plan(tape, params)
    names, A = P.args()
    M = np.zeros(runs.FRONT + P.mem_words + 8 * tape["fieldStride"] // 4, dtype=np.uint32)
    view = M[runs.FRONT:]
    view[: P.mem_words] = P.memory(str(tmp_path / "state"))
    prog = ol.Program(names, A)
    kw = dict(conf, input_sounding=open(snd).read(), sounding_winds=1) if snd else dict(conf)
    ref = runs.OracleRun(te_host.config(fuse=0, **kw))
    arena = P.base[len(tape["buffers"]) - 1]
    words = len(tape["fields"]) * tape["fieldStride"] // 4
    assert np.array_equal(view[arena: arena + words], ref.M[runs.arena_end(ref.enc) - words: runs.arena_end(ref.enc)])
    for _ in range(2):
        prog.run(view)
        ref.step(1)
        assert np.array_equal(view[arena: arena + words],
                              ref.M[runs.arena_end(ref.enc) - words: runs.arena_end(ref.enc)])


@pytest.mark.gpu
def test_replay_supercell_tape_on_gpu(gpu, tmp_path):
    out = tmp_path / "hashes.bin"
    print(node("replay", out, timeout=600), flush=True)
    got = np.fromfile(out, dtype=np.uint32)
    # the slots the oracle reproduces must all match on the GPU through the JS host
    P = golden_check.plan()
    have = set(ol.kernel_names()) | {"_copy"}
    names, A = P.args()
    prog = ol.Program(names, A)
    M = P.memory()
    ok_slots = []
    ok, bad, skipped, mism = golden_check.check(
        P, lambda i: prog.run(M, i, i), lambda off, n: ol.region_hash(M, off, n), have, ok_slots)
    assert bad == 0 and ok >= 800
    ref = tapeplan.ref_hashes()
    wrong = [s for s in ok_slots if got[s] != ref[s]]
    assert not wrong, wrong[:10]


@pytest.mark.gpu
@pytest.mark.parametrize("tune,sounding", [("", ""), ("tune", ""), ("", "hurricane")])
def test_js_run_bit_exact_vs_oracle(gpu, tune, sounding):
    """tune: run.autotune's launch shapes (forced) leave the fields as the
    oracle's; hurricane: buildRun(conf, sounding) with the reference's
    hurricane_3d input_sounding steps as the oracle does."""
    conf = dict(nx=32, ny=24, nz=16, tke=1, initial_tke=3, rayleigh=1)
    kw = dict(conf)
    args = ["run", json.dumps(conf), 2, tune]
    if sounding:
        path = os.path.join(HERE, "golden", "soundings", "hurricane_3d.input_sounding")
        args.append(path)
        kw.update(input_sounding=open(path).read(), sounding_winds=1)
    got = json.loads(node(*args))
    ref = runs.OracleRun(te_host.config(**kw))
    ref.step(2)
    e = ref.enc
    n = e.sJ * e.sK * (e.nz + 1)
    bad = [f for f in e.fields if f not in runs.FUSED_SCRATCH
           and got[f] != fnv(ref.M[e.field_offset(f): e.field_offset(f) + n])]
    assert not bad, bad


@pytest.mark.gpu
def test_js_adjoint_gradient_matches_the_c_abi(gpu, tmp_path):
    """LargeStepFunction (forward / backward) gives the same gradient of
    sum w^2 over a box w.r.t. the initial th, u, v as te_host.Adjoint."""
    conf = dict(nx=20, ny=16, nz=10, dz=2000.0)
    out = tmp_path / "grad.bin"
    node("adjoint", json.dumps(conf), 2, out, timeout=600)
    js = np.fromfile(out, dtype=np.float64)
    run = te_host.Run(te_host.config(**conf))
    run.init_state()
    adj = te_host.Adjoint(run, every=2)
    adj.forward(2)
    end = run.field_offset(run.fields[0]) + run.field_words * len(run.fields)
    state = run.read(0, end).view(np.float32)
    box = te_host.Box(run, i=(7, 13), j=(5, 11), k=(1, 6))
    seed = np.zeros(end, np.float32)
    seed[box.words] = 2.0 * state[box.words]
    adj.upload(0, seed)
    adj.backward()
    py = te_host.Controls(run, ("th", "u", "v")).pull(adj.read(0, end))
    adj.close()
    run.close()
    assert js.shape == py.shape
    assert np.linalg.norm(js - py) <= 1e-5 * np.linalg.norm(py)
