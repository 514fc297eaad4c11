"""Reference-tape planning for the tests (no GPU, no torch).

Restates the memory planning of the reference replayer (web/replay.js:46-121,
168-178 of ThunderEngine): every buffer of a tape gets a word range of one
flat memory M aligned to 64 words, each setBytes payload gets words after the
buffers, and each dispatch becomes a 64-word argument slot
(A[0..2] = thread grid, A[3] = 0, A[4+n] = word offset of parameter n).
Region-hash slots are assigned per (dispatch, written parameter) in the order
of kernels.json, exactly as replay.js:111-115 assigns them.
"""
import base64
import gzip
import json
import os

import numpy as np

ALIGN = 64
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
SUPERCELL = os.path.join(GOLDEN, "supercell")
# recorded presets with a fixture under tests/golden/<name>/
PRESETS = ("supercell", "squall_line", "supercell_stretched")


def load_preset(name):
    d = os.path.join(GOLDEN, name)
    with gzip.open(os.path.join(d, "tape_step1.json.gz"), "rt") as fh:
        tape = json.load(fh)
    params = json.load(open(os.path.join(d, "kernel_params.json")))
    tape["_dir"] = d
    return tape, params


def load_supercell():
    return load_preset("supercell")


class Plan:
    def __init__(self, tape, params, cmds=None):
        self.tape = tape
        self.params = params
        cmds = tape["steps"][0] if cmds is None else cmds
        top = 0
        self.base = []
        for b in tape["buffers"]:
            self.base.append(top)
            top += -(-b["length"] // (4 * ALIGN)) * ALIGN
        self.buffers_end = top
        self.bytes = []                       # (word offset, uint32 words)
        offs = {}
        for i, c in enumerate(cmds):
            if c["op"] != "bytes":
                continue
            raw = base64.b64decode(c["data"])
            words = -(-len(raw) // 4)
            pad = raw + b"\0" * (words * 4 - len(raw))
            offs[i] = top
            self.bytes.append((top, np.frombuffer(pad, dtype=np.uint32).copy()))
            top += -(-words // 4) * 4 + 4
        self.mem_words = -(-top // ALIGN) * ALIGN
        self.ops = []                         # dicts: name, grid, tg, bind | _copy n, src, dst
        self.slots = []                       # (op index, param index, word offset, length)
        bind, last, pipe = {}, {}, None
        for i, c in enumerate(cmds):
            op = c["op"]
            if op == "pipe":
                pipe = c["name"]
            elif op == "buf":
                bind[c["index"]] = 0 if c["buf"] < 0 else self.base[c["buf"]] + c["offset"] // 4
                last[c["index"]] = c["buf"]
            elif op == "bufoff":
                bind[c["index"]] = self.base[last[c["index"]]] + c["offset"] // 4
            elif op == "bytes":
                bind[c["index"]] = offs[i]
            elif op == "end":
                bind = {}
            elif op == "copy":
                self.ops.append({"name": "_copy", "n": c["size"] // 4,
                                 "src": self.base[c["src"]] + c["srcOffset"] // 4,
                                 "dst": self.base[c["dst"]] + c["dstOffset"] // 4})
            elif op == "dispatch":
                o = {"name": pipe, "grid": list(c["grid"]), "tg": list(c["tg"]), "bind": dict(bind)}
                for idx, kind, write in params[pipe]:
                    if write:
                        off = bind.get(idx, 0)
                        self.slots.append((len(self.ops), idx, off, self.region_len(off)))
                self.ops.append(o)

    def region_len(self, off):
        """Words a kernel writes for a parameter: the arena field or the rest of the buffer."""
        t = self.tape
        for b in range(len(t["buffers"]) - 1, -1, -1):
            if off >= self.base[b]:
                end = self.base[b] + t["buffers"][b]["length"] // 4
                if t["buffers"][b]["arena"]:
                    return min(t["fieldStride"] // 4, end - off)
                return end - off
        return 0

    def args(self):
        """names, (nops, 64) uint32 argument slots (copies: A[0]=n, A[4]=src, A[5]=dst)."""
        names, A = [], np.zeros((len(self.ops), 64), dtype=np.uint32)
        for n, o in enumerate(self.ops):
            names.append(o["name"])
            if o["name"] == "_copy":
                A[n, 0], A[n, 1], A[n, 2] = o["n"], 1, 1
                A[n, 4], A[n, 5] = o["src"], o["dst"]
                continue
            A[n, 0:3] = o["grid"]
            for idx, off in o["bind"].items():
                A[n, 4 + idx] = off
        return names, A

    def memory(self, state_dir=None):
        """Flat u32 memory with the t=0 buffers and the parameter bytes in place."""
        M = np.zeros(self.mem_words, dtype=np.uint32)
        state_dir = state_dir or os.path.join(self.tape.get("_dir", SUPERCELL), "state")
        for b in self.tape["buffers"]:
            p = os.path.join(state_dir, "%d.bin.gz" % b["id"])
            if os.path.exists(p):
                raw = gzip.open(p, "rb").read()
                w = np.frombuffer(raw[: len(raw) // 4 * 4], dtype=np.uint32)
                M[self.base[b["id"]]: self.base[b["id"]] + len(w)] = w
        for off, w in self.bytes:
            M[off: off + len(w)] = w
        return M

    def field_of(self, off):
        """Arena field name (or buffer label) that contains word offset `off`."""
        t = self.tape
        for b in range(len(t["buffers"]) - 1, -1, -1):
            if off >= self.base[b]:
                if t["buffers"][b]["arena"]:
                    rel = (off - self.base[b]) * 4
                    best = None
                    for name, o in t["fields"].items():
                        if o <= rel and (best is None or o > t["fields"][best]):
                            best = name
                    return best
                return t["buffers"][b]["label"] or ("buf%d" % b)
        return None


def ref_hashes(name="supercell"):
    return np.fromfile(os.path.join(GOLDEN, name, "hashes.ref.bin"), dtype=np.uint32)
