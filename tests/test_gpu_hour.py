"""The supercell hour on the MI355X: 600 large steps of the 192x192x80
benchmark workload (fused launches on autotuned shapes, graph replays) give
the oracle's bits in every field the step carries.

The expected values are the word-wise FNV-1a hashes (web/view.js:1064-1068)
of each FieldArena field and of the profile words after steps 1, 10, 100 and
600 of the oracle stepping the reference's dispatch list
(tests/golden/hour_192x192x80.json, written by
tests/golden/make_hour_hashes.py in about 20 minutes on 8 cores).
"""
import json
import os
import sys

import pytest

import bench
import runs
import te_host

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_supercell_hour_field_hashes(gpu):
    data = json.load(open(os.path.join(HERE, "golden", "hour_192x192x80.json")))
    assert data["workload"] == bench.WORKLOAD
    assert "600" in data["checkpoints"]
    run = te_host.Run(te_host.config(**bench.WORKLOAD))
    run.init_state()
    say("dispatches on tuned shapes", run.autotune(reps=2, force=True))
    done = 0
    for n in sorted(int(s) for s in data["checkpoints"]):
        run.step(n - done)
        run.sync()
        done = n
        bad = []
        for f, want in data["checkpoints"][str(n)].items():
            if f in runs.FUSED_SCRATCH:  # dead scratch: fused launches may leave it different
                continue
            if f == "_profiles":
                got = run.fnv1a(0, run.field_offset(run.fields[0]))
            else:
                got = run.fnv1a(run.field_offset(f), run.field_words)
            if got != want:
                bad.append(f)
        say("after step", n, "fields", len(data["checkpoints"][str(n)]), "differing", bad)
        assert not bad, (n, bad)
    run.close()
