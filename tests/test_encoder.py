"""The C++ encoder (te_encode) against the reference's recorded tapes (CPU).

The supercell recordings hold the reference encoder's output for step 1
(web/core/timestep.js TimeStep.encode) of the 120x120x40 supercell and of the
same storm on the 144x144x40 stretched grid (500 m over the central 40 km
growing to 2 km at +-60 km, data/presets/supercell_stretched.json).
te_encode with the same options must produce the same dispatch sequence
(the stretched grid's acoustic substep count follows from its 500 m finest
spacing), launch grids and parameter bindings (arena field + word offset,
profile/parameter buffer + offset, and the content of every setBytes
payload) for every dispatch of the dynamics path, i.e. up to the
microphysics tail, whose kernels are not on this path; and the same t=0
content of every profile and parameter buffer the recording holds
(hmetric: the stretched grid's map factors and face positions).
"""
import gzip
import os

import numpy as np
import pytest

import golden_check
import te_host

TERRAIN = range(20, 30)  # terrain bindings (always the zero terrain buffer)
REGIONS = ["c1", "c2", "mh", "mf", "zf", "rho0", "rr0", "rf0", "th0", "pi0", "rth0", "qv0", "qc0", "qi0",
           "tauh", "tauf", "u0", "v0", "hmetric", "radbcw", "radbce", "radbcs", "radbcn", "sfc",
           "mass_rows", "mass_levels", "mass_ref", "mass_out"]
MASS_LABEL = {"mass_rows": "buf30", "mass_levels": "buf31", "mass_ref": "buf32", "mass_out": "buf33"}


HERE = os.path.dirname(os.path.abspath(__file__))
# data/presets/supercell_stretched.json: stretch_x = stretch_y = 1, dx_inner 500,
# dx_outer 2000, nos_x_len 40 km, tot_x_len 120 km (and the same in y)
STRETCHED = dict(nx=144, ny=144, dx=1250., dy=1250., dt=3., stretch_x=1, stretch_y=1, dx_inner=500.,
                 dx_outer=2000., nos_x_len=40000., tot_x_len=120000., dy_inner=500., dy_outer=2000.,
                 nos_y_len=40000., tot_y_len=120000.)
PRESETS = {"supercell": {}, "supercell_stretched": STRETCHED}


def supercell_config(fields, **kw):
    base = dict(nx=120, ny=120, nz=40, dx=1000., dy=1000., dz=500., dt=6., nsound=0, moist=1, species=11,
                open_x=1, open_y=1, weno_scalars=1, weno_momentum=0, tke=1, explicit_vdiff=0,
                dissipative_heating=1, rayleigh=1, rayleigh_zd=15000., umove=12.5, vmove=3.0, microphysics=5,
                mass_conservation=1, cache_moisture_turbulence=1, arena_fields=fields, fuse=0)
    base.update(kw)
    return te_host.config(**base)


@pytest.fixture(scope="module", params=sorted(PRESETS))
def encoded(request):
    P = golden_check.plan(request.param)
    fields = [n for n, _ in sorted(P.tape["fields"].items(), key=lambda x: x[1])]
    r = te_host.Run(supercell_config(fields, **PRESETS[request.param]), encode_only=True)
    names, A = r.program()
    return P, fields, r, names, A


@pytest.mark.parametrize("preset", sorted(PRESETS))
def test_profile_buffers_match_recording(preset):
    """Every recorded t=0 profile / parameter buffer of the dynamics path
    (state/<id>.bin.gz) equals te_encode's region word for word; hmetric
    carries the horizontal grid (map factors uh uf vh vf, faces, centres)."""
    P = golden_check.plan(preset)
    fields = [n for n, _ in sorted(P.tape["fields"].items(), key=lambda x: x[1])]
    r = te_host.Run(supercell_config(fields, **PRESETS[preset]), encode_only=True)
    compared = []
    for b in P.tape["buffers"]:
        name = b["label"]
        if b["arena"] or name not in ("hmetric", "c1", "c2", "mh", "mf", "rf0", "rr0", "rho0", "tauh", "tauf"):
            continue
        path = os.path.join(HERE, "golden", preset, "state", "%d.bin.gz" % b["id"])
        want = np.frombuffer(gzip.open(path).read(), dtype=np.uint32)
        got = r.read(r.region_offset(name), want.size)
        assert np.array_equal(got, want), (preset, name, np.flatnonzero(got != want)[:8])
        compared.append(name)
    assert "hmetric" in compared, compared


def test_dispatch_sequence(encoded):
    P, fields, r, names, A = encoded
    ref = [o["name"] for o in P.ops]
    mi = ref.index("mp_pdefq")
    assert names.index("mp_pdefq") == mi
    assert names[:mi] == ref[:mi]
    for i in range(mi):
        assert list(A[i, 0:3]) == list(P.ops[i]["grid"]), (i, names[i])


def test_bindings(encoded):
    P, fields, r, names, A = encoded
    arena0, fw = r.field_offset(fields[0]), r.field_words
    regions = {}
    for n in REGIONS:
        try:
            regions[n] = r.region_offset(n)
        except te_host.TeError:
            pass
    starts = sorted((o, n) for n, o in regions.items())
    img = r.read(0, r.mem_words)

    def mine(off):
        if arena0 <= off < arena0 + fw * len(fields):
            q = (off - arena0) // fw
            return ("field", fields[q], off - arena0 - q * fw)
        if off >= arena0 + fw * len(fields):  # setBytes payloads follow the arena
            return ("bytes", off)
        cand = [(o, n) for o, n in starts if o <= off]
        if cand:
            o, n = cand[-1]
            if n == "sfc" or n in MASS_LABEL or off == o:
                return ("buf", MASS_LABEL.get(n, n), off - o)
        return ("bytes", off)

    def ref(off):
        if off >= P.buffers_end:
            return ("bytes", [w for o, w in P.bytes if o == off][0])
        for b in range(len(P.tape["buffers"]) - 1, -1, -1):
            if off >= P.base[b]:
                if P.tape["buffers"][b]["arena"]:
                    f = P.field_of(off)
                    return ("field", f, off - P.base[b] - P.tape["fields"][f] // 4)
                return ("buf", P.tape["buffers"][b]["label"] or "buf%d" % b, off - P.base[b])

    mi = [o["name"] for o in P.ops].index("mp_pdefq")
    bad = []
    checked = 0
    for i in range(mi):
        o = P.ops[i]
        if o["name"] == "_copy":
            continue
        for idx, kind, write in P.params[o["name"]]:
            if idx in TERRAIN or idx not in o["bind"]:
                continue
            rd, md = ref(o["bind"][idx]), mine(int(A[i, 4 + idx]))
            if rd[0] == "bytes":
                w = rd[1]
                got = img[md[1]: md[1] + len(w)] if md[0] == "bytes" else None
                if got is None or not np.array_equal(got, w):
                    bad.append((i, o["name"], idx, "bytes", list(w[:4]), None if got is None else list(got[:4])))
            elif rd != md:
                bad.append((i, o["name"], idx, rd, md))
            checked += 1
    assert checked > 5000
    assert not bad, bad[:10]


def test_schedule_closed_form():
    """Acoustic substeps per large step: nsound n -> stages of n1, n/2, n
    substeps with n1 = max(1, round(n/3)) (timestep.js schedule), i.e. 11 for
    n = 6; dispatch counts follow it."""
    for ns, total in ((6, 2 + 3 + 6), (4, 2 + 2 + 4), (8, 3 + 4 + 8)):
        r = te_host.Run(te_host.config(nx=16, ny=16, nz=12, nsound=ns, fuse=0), encode_only=True)
        names, _ = r.program()
        assert names.count("acoustic_uv") == total, (ns, names.count("acoustic_uv"))
        assert names.count("acoustic_w_solve") == total
        r = te_host.Run(te_host.config(nx=16, ny=16, nz=12, nsound=ns, fuse=2), encode_only=True)
        names, _ = r.program()
        assert names.count("acoustic_uv") == names.count("acoustic_w_solve_lds") == names.count("ac_pp_accumulate") == total
        r = te_host.Run(te_host.config(nx=16, ny=16, nz=12, nsound=ns, fuse=1), encode_only=True)
        names, _ = r.program()
        assert names.count("acoustic_uv") == names.count("ac_column") == total
        assert "acoustic_prepare" not in names and "acoustic_begin" not in names


def test_encoder_rejects_bad_config():
    with pytest.raises(te_host.TeError):
        te_host.Run(te_host.config(nx=0), encode_only=True)
    with pytest.raises(te_host.TeError):
        te_host.Run(te_host.config(species=5), encode_only=True)
    with pytest.raises(te_host.TeError):  # automatic schedule beyond 192 substeps
        te_host.Run(te_host.config(nsound=0, dt=300.0, dx=100.0, dy=100.0), encode_only=True)


HAZARD_CONFIGS = [
    dict(nx=24, ny=20, nz=12, tke=1, rayleigh=1),
    dict(nx=24, ny=20, nz=12, tke=1, species=7, weno_momentum=1),
    dict(nx=24, ny=20, nz=12, tke=0, moist=0, rayleigh=1),
    dict(nx=24, ny=20, nz=12, tke=1, open_x=1, open_y=1, mass_conservation=1),
]


@pytest.mark.parametrize("kw", HAZARD_CONFIGS, ids=lambda kw: "-".join(f"{k}{v}" for k, v in kw.items()))
def test_fused_launches_write_no_field_their_neighbours_read(kw):
    """A launch whose workgroups read a field at neighbouring points must not
    write that field: workgroups run in no fixed order, so the reads could see
    the new values.  The oracle runs a fused launch as the reference dispatches
    it replaces, one after another, and cannot show such a race; this checks
    the bindings of every fused launch instead."""
    r = te_host.Run(te_host.config(fuse=1, **kw), encode_only=True)
    names, A = r.program()

    def word(op, idx):
        return int(r.read(int(A[op, 4 + idx]), 1).view(np.int32)[0])

    seen = set()
    for op, name in enumerate(names):
        off = lambda idx: int(A[op, 4 + idx])  # noqa: E731
        if name == "adv_fused":
            stencil = {off(2), off(4), off(13), off(14)}
            chain = word(op, 32)
            written = {off(40)} if chain & 8 else {off(11)}
            assert not written & stencil, (op, name, chain)
            if chain & 8:  # the chained update writes the target, reads its start-of-step copy
                assert off(39) != off(40)
            seen.add(name)
        elif name == "ac_pp_accumulate":  # w is read at k and k+1
            written = {off(q) for q in (7, 8, 9, 13, 14, 15)}
            assert off(2) not in written, (op, name)
            seen.add(name)
        elif name == "ac_column":
            # one column per thread: w, pp3d are read and written in the own
            # column only; piadv is read at neighbouring columns and never
            # written; u, v are read at i+1 / j+1 and written only at halo
            # cells, by the one thread that reads them there
            written = {off(q) for q in (4, 10, 11, 12, 13, 14, 15)}
            assert off(5) not in written and not {off(q) for q in (6, 7, 8, 9)} & written, (op, name)
            assert off(2) not in written and off(3) not in written, (op, name)
            seen.add(name)
        elif name == "massflux_div":  # u, v, w read at i+1, j+1, k+1; the fluxes and their divergence written
            written = {off(q) for q in (6, 7, 8, 10)}
            assert not {off(1), off(2), off(3)} & written, (op, name)
            seen.add(name)
        elif name == "pdef_dum_xy":  # flux and mass flux read at the next face; the output and its images written
            assert off(10) not in {off(2), off(3), off(4), off(5)}, (op, name)
            seen.add(name)
        elif name == "add_fields":  # independent triples
            n = word(op, 13)
            srcs, dsts = {off(1 + q) for q in range(n)}, [off(5 + q) for q in range(n)]
            assert len(set(dsts)) == n and not srcs & set(dsts), (op, name)
            seen.add(name)
        elif name == "copy_fields":
            n = word(op, 9)
            srcs, dsts = {off(1 + q) for q in range(n)}, [off(5 + q) for q in range(n)]
            assert len(set(dsts)) == n and not srcs & set(dsts), (op, name)
            seen.add(name)
    assert "adv_fused" in seen and ("ac_pp_accumulate" in seen or "ac_column" in seen)
