"""Writes kernel_io.json: per kernel, its buffer parameters as
[index, read, written] (bench.py's algorithmic-byte count per launch).

`written` is the write flag of the reference's kernel signature table
(data/kernels/kernels.json, carried in tests/golden/supercell/kernel_params.json);
`read` is true for parameters that are not written and for written ones the
HIP kernel also loads (read-modify-write), found in csrc/k_*.hip.
"""
import glob
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def main():
    params = json.load(open(os.path.join(ROOT, "tests", "golden", "supercell", "kernel_params.json")))
    src = "".join(open(f).read() for f in sorted(glob.glob(os.path.join(HERE, "csrc", "k_*.hip"))))
    bodies = {}
    for part in re.split(r"__global__ void k_", src)[1:]:
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
                m = re.search(r"\*?(\w+)\s*=\s*fp\(M,\s*a\.p\[%d\]\)" % idx, body)
                if m:
                    nm = m.group(1)
                    uses = len(re.findall(r"\b%s\[" % nm, body))
                    stores = len(re.findall(r"\b%s\[[^\]]*\]\s*(?:[-+*/]?=)(?!=)" % nm, body))
                    compound = len(re.findall(r"\b%s\[[^\]]*\]\s*[-+*/]=" % nm, body))
                    read = uses > stores or compound > 0
                else:
                    direct = re.findall(r"fp\(M,\s*a\.p\[%d\]\)\[[^\]]*\]\s*(=|[-+*/]=)?" % idx, body)
                    read = any(op != "=" for op in direct)
            rows.append([idx, bool(read), bool(w)])
        out[k] = rows
    json.dump(out, open(os.path.join(HERE, "kernel_io.json"), "w"), indent=0, sort_keys=True)


if __name__ == "__main__":
    main()
