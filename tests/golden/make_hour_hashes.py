"""Writes the oracle's per-field FNV-1a hashes of the supercell hour.

Workload: the 192x192x80 benchmark configuration (bench.WORKLOAD: dx = dy =
1 km, dz = 250 m, dt = 6 s, nsound = 6, TKE, Rayleigh damping, WENO5 scalars
th qv qc, synthetic WK82 + bubble t=0 state), stepped by the oracle along the
reference's dispatch list (fuse = 0) for 600 large steps (one model hour).
After steps 1, 10, 100 and 600 it records the word-wise FNV-1a
(web/view.js:1064-1068) of every FieldArena field the step carries (all but
the scratch fields fused launches may leave different) and of the profile /
parameter words below the arena.

    python tests/golden/make_hour_hashes.py [steps]

takes about 17 minutes on 8 cores.  The output (hour_192x192x80.json) is data:
expected values for tests/test_gpu_hour.py, which steps the HIP path the same
number of steps and compares te_fnv1a of each field on the device.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "thunderengine-web_amd")):
    sys.path.insert(0, p)

import bench  # noqa: E402
import oracle_lib  # noqa: E402
import runs  # noqa: E402
import te_host  # noqa: E402

CHECKPOINTS = (1, 10, 100, 600)
OUT = os.path.join(HERE, "hour_192x192x80.json")


def hashes(R):
    e = R.enc
    out = {"_profiles": oracle_lib.fnv1a(R.M[: e.field_offset(e.fields[0])])}
    for f in e.fields:
        if f in runs.FUSED_SCRATCH:
            continue
        o = e.field_offset(f)
        out[f] = oracle_lib.fnv1a(R.M[o: o + e.field_words])
    return out


def main():
    last = int(sys.argv[1]) if len(sys.argv) > 1 else CHECKPOINTS[-1]
    cfg = te_host.config(**bench.WORKLOAD)
    R = runs.OracleRun(cfg)
    res = {"workload": bench.WORKLOAD, "dispatches_per_step": len(R.names), "checkpoints": {}}
    t0 = time.time()
    for n in range(1, last + 1):
        R.step(1)
        if n in CHECKPOINTS or n == last:
            res["checkpoints"][str(n)] = hashes(R)
            print("step %d: %.0f s" % (n, time.time() - t0), flush=True)
            with open(OUT, "w") as fh:
                json.dump(res, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
