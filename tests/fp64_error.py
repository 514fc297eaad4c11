"""Error of the FP32 large step against the FP64 build of the same step.

Runs the oracle's float and double instantiations of one dispatch list from the
same t=0 state and reports, for u, v, w, theta' and pi' over the interior,
the largest absolute difference and the RMS difference after each requested
step count.  Because the HIP step is bit-identical to the FP32 oracle, these
are the HIP step's errors against FP64.

    python tests/fp64_error.py [--nx 64 --ny 64 --nz 40] [--steps 1,600]
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.join(os.path.dirname(HERE), "thunderengine-web_amd")]

import runs  # noqa: E402
import te_host  # noqa: E402

FIELDS = {"u": "u", "v": "v", "w": "w", "theta'": "th", "pi'": "pp3d"}


def run(cfg, checkpoints):
    r32 = runs.OracleRun(cfg)
    e = r32.enc
    M64 = r32.M.copy()
    V = r32.M.view(np.float32).astype(np.float64)
    n = e.sJ * e.sK * (e.nz + 1)
    out, done = {}, 0
    for target in sorted(checkpoints):
        for _ in range(target - done):
            r32.prog.run(r32.M)
            r32.prog.run_f64(M64, V)
        done = target
        row = {}
        for label, f in FIELDS.items():
            o = e.field_offset(f)
            a = r32.M[o: o + n].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)[: e.nz, 3: 3 + e.ny, 3: 3 + e.nx]
            b = V[o: o + n].reshape(e.nz + 1, e.sK, e.sJ)[: e.nz, 3: 3 + e.ny, 3: 3 + e.nx]
            d = a.astype(np.float64) - b
            row[label] = {"max_abs": float(np.abs(d).max()), "rms": float(np.sqrt(np.mean(d * d))),
                          "max_abs_fp64": float(np.abs(b).max())}
        out[target] = row
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--dz", type=float, default=500.0)
    ap.add_argument("--tke", type=int, default=0)
    ap.add_argument("--steps", default="1,600")
    a = ap.parse_args()
    cfg = te_host.config(nx=a.nx, ny=a.ny, nz=a.nz, dz=a.dz, tke=a.tke, rayleigh=1 if a.tke else 0)
    res = run(cfg, [int(s) for s in a.steps.split(",")])
    print(json.dumps({"config": vars(a), "errors": res}, indent=1))


if __name__ == "__main__":
    main()
