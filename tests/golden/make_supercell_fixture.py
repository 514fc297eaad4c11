"""Writes the supercell golden fixture from the reference checkout.

Run in the build container (the reference tree is not present on GPU hosts):

    python tests/golden/make_supercell_fixture.py [/root/reference]

It copies, unchanged, the reference's own recording of one supercell step
(data/supercell/ of ThunderEngine): the t=0 content of every buffer
(state/<id>.bin.gz), the per-kernel region hashes of step 1 (hashes.ref.bin),
the per-buffer FNV-1a after step 1 (step1.json) and grid.json.  From tape.json
it keeps the header and the command list of step 1 only (gzip'd JSON), and from
data/kernels/kernels.json the parameter tables of the kernels that list uses.
These are data files (inputs and expected outputs), AGPL-3.0, see
tests/golden/supercell/PROVENANCE.txt.
"""
import gzip
import json
import os
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "data", "supercell")
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "supercell")
os.makedirs(os.path.join(dst, "state"), exist_ok=True)

for f in os.listdir(os.path.join(src, "state")):
    shutil.copyfile(os.path.join(src, "state", f), os.path.join(dst, "state", f))
for f in ("hashes.ref.bin", "step1.json", "grid.json"):
    shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))

tape = json.load(open(os.path.join(src, "tape.json")))
step = tape["steps"][0]
head = {k: v for k, v in tape.items() if k not in ("steps", "diag")}
head["steps"] = [step]
with gzip.open(os.path.join(dst, "tape_step1.json.gz"), "wt") as fh:
    json.dump(head, fh, separators=(",", ":"))

kj = json.load(open(os.path.join(ref, "data", "kernels", "kernels.json")))
used = sorted({c["name"] for c in step if c["op"] == "pipe"})
params = {n: [[p["index"], p["kind"], int(p["write"])] for p in kj[n]["params"]] for n in used}
with open(os.path.join(dst, "kernel_params.json"), "w") as fh:
    json.dump(params, fh, indent=0, sort_keys=True)
print("wrote", dst, "kernels:", len(used))
