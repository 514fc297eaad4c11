"""Writes kernel_io.json: per kernel, its buffer parameters as
[index, read, written] (bench.py's algorithmic-byte count per launch).

`written` is the write flag of the reference's kernel signature table
(data/kernels/kernels.json, carried in tests/golden/supercell/kernel_params.json).
`read` is true for parameters that are not written and for written
ones the HIP kernel also loads (read-modify-write), found in csrc/k_*.hip.
Kernels of this path missing from that table are listed in EXTRA, from their
WGSL bindings.
"""
import glob
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


# reference kernels on this path absent from the recorded signature table:
# [index, kind, written] as in kernels.json ("ptr" storage, "ref" uniform)
EXTRA = {
    # update_diagnostics.wgsl (dry): th, pp, th0, pi0 -> thv, ppterm, piadv; rth0; buoy
    "update_diagnostics": [[0, "ref", 0], [1, "ptr", 0], [2, "ptr", 0], [3, "ptr", 0], [4, "ptr", 0],
                           [5, "ptr", 1], [6, "ptr", 1], [7, "ptr", 1], [8, "ptr", 0], [9, "ptr", 1]],
    # adv_flux_{x,y}_weno_uv.wgsl: stag, s, rr, weps -> flux
    "adv_flux_x_weno_uv": [[0, "ref", 0], [1, "ref", 0], [2, "ptr", 0], [3, "ptr", 0], [4, "ref", 0], [5, "ptr", 1]],
    "adv_flux_y_weno_uv": [[0, "ref", 0], [1, "ref", 0], [2, "ptr", 0], [3, "ptr", 0], [4, "ref", 0], [5, "ptr", 1]],
    # adv_flux_z_weno_uv.wgsl: stag, s, rrw, c1, c2, weps -> fz
    "adv_flux_z_weno_uv": [[0, "ref", 0], [1, "ref", 0], [2, "ptr", 0], [3, "ptr", 0], [4, "ptr", 0], [5, "ptr", 0],
                           [6, "ref", 0], [7, "ptr", 1]],
    # bc_periodic_{x,y}.wgsl: field (read-write), pass
    "bc_periodic_x": [[0, "ref", 0], [1, "ptr", 1], [2, "ref", 0]],
    "bc_periodic_y": [[0, "ref", 0], [1, "ptr", 1], [2, "ref", 0]],
    # copy_words (replay.js:403-406): source -> destination
    "_copy": [[0, "ptr", 0], [1, "ptr", 1]],
}


def main():
    params = json.load(open(os.path.join(ROOT, "tests", "golden", "supercell", "kernel_params.json")))
    params = {k: v for k, v in params.items() if not k.startswith(("morr_", "mp_"))}
    params.update(EXTRA)
    src = "".join(open(f).read() for f in sorted(glob.glob(os.path.join(HERE, "csrc", "k_*.hip"))))
    bodies = {}
    for part in re.split(r"__global__ (?:__launch_bounds__\(\d+\) )?void k_", src)[1:]:
        bodies[part.split("(")[0]] = part.split("\n}\n")[0]
    out = {}
    for k, plist in sorted(params.items()):
        body = bodies.get(k, "")
        rows = []
        for idx, kind, w in plist:
            if kind != "ptr":
                continue
            read = not w
            if w and body:
                m = re.search(r"\*?(\w+)\s*=\s*fq\(a\.q\[%d\]\)" % idx, body)
                if m:
                    nm = m.group(1)
                    uses = len(re.findall(r"\b%s\[" % nm, body))
                    stores = len(re.findall(r"\b%s\[[^\]]*\]\s*(?:[-+*/]?=)(?!=)" % nm, body))
                    compound = len(re.findall(r"\b%s\[[^\]]*\]\s*[-+*/]=" % nm, body))
                    read = uses > stores or compound > 0
                else:
                    direct = re.findall(r"fq\(a\.q\[%d\]\)\[[^\]]*\]\s*(=|[-+*/]=)?" % idx, body)
                    read = any(op != "=" for op in direct)
            rows.append([idx, bool(read), bool(w)])
        out[k] = rows
    json.dump(out, open(os.path.join(HERE, "kernel_io.json"), "w"), indent=0, sort_keys=True)


if __name__ == "__main__":
    main()
