"""Writes a golden fixture (one recorded preset step) from the reference checkout.

Run in the build container (the reference tree is not present on GPU hosts):

    python tests/golden/make_fixture.py <preset> [/root/reference]

with <preset> one of the reference's recordings (supercell, squall_line,
supercell_stretched, ...).  It copies, unchanged, the reference's own
recording of one step (data/<preset>/ of ThunderEngine): the t=0 content of every buffer
(state/<id>.bin.gz), the per-kernel region hashes of step 1 (hashes.ref.bin),
the per-buffer FNV-1a after step 1 (step1.json) and grid.json.  From tape.json
it keeps the header and the command list of step 1 only (gzip'd JSON), and from
data/kernels/kernels.json the parameter tables of the kernels that list uses.
These are data files (inputs and expected outputs), AGPL-3.0, see
tests/golden/<preset>/PROVENANCE.txt.
"""
import gzip
import json
import os
import shutil
import sys

preset = sys.argv[1]
ref = sys.argv[2] if len(sys.argv) > 2 else "/root/reference"
src = os.path.join(ref, "data", preset)
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), preset)
os.makedirs(os.path.join(dst, "state"), exist_ok=True)

for f in os.listdir(os.path.join(src, "state")):
    shutil.copyfile(os.path.join(src, "state", f), os.path.join(dst, "state", f))
for f in ("hashes.ref.bin", "step1.json", "grid.json"):
    shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))

tape = json.load(open(os.path.join(src, "tape.json")))
step = tape["steps"][0]
head = {k: v for k, v in tape.items() if k not in ("steps", "diag")}
head["steps"] = [step]
with gzip.GzipFile(os.path.join(dst, "tape_step1.json.gz"), "wb", mtime=0) as fh:
    fh.write(json.dumps(head, separators=(",", ":")).encode())

kj = json.load(open(os.path.join(ref, "data", "kernels", "kernels.json")))
used = sorted({c["name"] for c in step if c["op"] == "pipe"})
params = {n: [[p["index"], p["kind"], int(p["write"])] for p in kj[n]["params"]] for n in used}
with open(os.path.join(dst, "kernel_params.json"), "w") as fh:
    json.dump(params, fh, indent=0, sort_keys=True)
nslots = os.path.getsize(os.path.join(src, "hashes.ref.bin")) // 4
ndisp = sum(1 for c in step if c["op"] == "dispatch")
with open(os.path.join(dst, "PROVENANCE.txt"), "w") as fh:
    fh.write(
        "Files in this folder are data recorded by ThunderEngine (defds7721-cyber/thunderengine-web),\n"
        "copyright (C) 2026 vvas2002, licensed AGPL-3.0-only; the model is CM1 r20.2 by George Bryan (NCAR).\n"
        "They come from data/%s/ and data/kernels/kernels.json of that repository and were\n"
        "extracted, unchanged apart from keeping only step 1 of tape.json and only the parameter\n"
        "tables of the kernels that step uses, by tests/golden/make_fixture.py.\n\n"
        "  state/<id>.bin.gz      t=0 content of buffer <id> of the %s preset (%dx%dx%d)\n"
        "  tape_step1.json.gz     tape header + command list of one large step (%d dispatches)\n"
        "  hashes.ref.bin         u32 region hash per (dispatch, written parameter) of step 1 (%d slots)\n"
        "  step1.json             FNV-1a of every buffer after step 1\n"
        "  grid.json              grid description of the preset\n"
        "  kernel_params.json     kernel -> [[param index, kind, write], ...]\n"
        % (preset, preset, tape["nx"], tape["ny"], tape["nz"], ndisp, nslots))
print("wrote", dst, "kernels:", len(used))
