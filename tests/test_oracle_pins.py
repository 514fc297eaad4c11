"""Pins of oracle functions that no reference recording covers (CPU).

No recorded preset uses momentum WENO (advwenov = 2), a dry atmosphere or a
stretched vertical grid, so these parts of the oracle are checked against
what the mathematics fixes, with checks that a plausible slip fails:

* momentum WENO5 (adv_flux_{x,y,z}_weno_uv, adv_flux_{x,y,z}_weno_w): on
  smooth fields WENO-Z's weights are the ideal ones, so every face value
  equals the 5th-order upwind flux flx5 of the flux-form kernels
  (adv_flux_*_stag, pinned by the supercell's recorded hashes) up to rounding,
  with mass fluxes of both signs (upwind side and stencil offsets); at a jump
  the nonlinear weights keep the face values inside the stencil's range where
  flx5 overshoots (adv_flux_x_weno_uv.wgsl:106-188,
  adv_flux_z_weno_uv.wgsl:106-236, adv_flux_x_weno_w.wgsl:143-175,
  adv_flux_z_weno_w.wgsl:159-231);
* the dry update_diagnostics (update_diagnostics.wgsl:115-144) equals the
  moist kernel (pinned by the recorded hashes) with every moisture field and
  profile zero: thv and piadv bit for bit, buoy within 2 ulp and ppterm
  within 4 ulp (the two kernels associate the same products differently and
  round rd/cv as rovcp * cp / cv against rd / (cp - rd));
* the stretched vertical grid (VerticalGrid(faceHeights), web/core/grid.js:
  256-293): face interpolation weights sum to one and interpolate linear
  profiles exactly, the metric factors telescope to the column depth, and the
  Fornberg surface / top extrapolation weights are exact on quadratics.
"""
import numpy as np
import pytest

import oracle_lib as ol
import runs
import te_host

GRID = dict(nx=40, ny=36, nz=24, dz=500.0)


def _program(kw):
    R = runs.OracleRun(te_host.config(fuse=0, **kw))
    return R, R.enc


def _field(M, e, name):
    o = e.field_offset(name)
    n = e.sJ * e.sK * (e.nz + 1)
    return M[o: o + n].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)


def _run_one(R, M, i):
    buf = np.zeros_like(R._buf)
    buf[runs.FRONT: runs.FRONT + M.size] = M
    R.prog.run(buf[runs.FRONT:], i, i)
    return buf[runs.FRONT: runs.FRONT + M.size]


# parameter index of the advected field and of the output flux per kernel
ADV_IO = {"adv_flux_x_stag": (2, 8), "adv_flux_y_stag": (2, 8), "adv_flux_z_stag": (2, 8),
          "adv_flux_x_weno_uv": (2, 5), "adv_flux_y_weno_uv": (2, 5), "adv_flux_z_weno_uv": (2, 7),
          "adv_flux_x_weno_w": (1, 5), "adv_flux_y_weno_w": (1, 5), "adv_flux_z_weno_w": (1, 4)}


def _pairs(R):
    """(weno op, flux-form op) pairs advecting the same field in the same
    direction: stage 3's WENO momentum fluxes and stage 1's flx5 fluxes."""
    names, A = R.names, R.args
    first = {}
    for i, n in enumerate(names):
        if n.endswith("_stag"):
            key = (n.split("_")[2], int(A[i, 4 + ADV_IO[n][0]]))
            first.setdefault(key, i)
    out = []
    for i, n in enumerate(names):
        if "_weno_uv" in n or "_weno_w" in n:
            key = (n.split("_")[2], int(A[i, 4 + ADV_IO[n][0]]))
            out.append((i, first[key]))
    return out


def _smooth_state(R, e, rng):
    """The t=0 memory with u, v, w, rru, rrv, rrw replaced by smooth fields
    (wavelengths of 128+ cells and 40+ levels) whose mass fluxes change sign
    inside the domain."""
    M = R.M0.copy()
    K, J, I = np.meshgrid(np.arange(e.nz + 1), np.arange(e.sK) - 3.0, np.arange(e.sJ) - 3.0, indexing="ij")
    for name in ("u", "v", "w", "rru", "rrv", "rrw"):
        p = rng.uniform(0, 2 * np.pi, 4)
        f = (8.0 * np.sin(2 * np.pi * I / 160.0 + p[0]) * np.cos(2 * np.pi * J / 140.0 + p[1])
             + 5.0 * np.sin(2 * np.pi * K / 48.0 + p[2]) + 2.0 * np.cos(2 * np.pi * (I + J) / 200.0 + p[3]))
        _field(M, e, name)[:] = f.astype(np.float32)
    return M


def _levels(name, e):
    # levels where both the WENO and the flux-form kernels use 5-point stencils
    if "_z_" in name:
        return slice(3, e.nz - 2)
    return slice(1, e.nz)


@pytest.mark.parametrize("seed", [1, 2])
def test_momentum_weno_equals_flx5_on_smooth_fields(seed):
    R, e = _program(dict(GRID, weno_momentum=1))
    pairs = _pairs(R)
    assert sorted({R.names[i] for i, _ in pairs}) == sorted(
        ["adv_flux_x_weno_uv", "adv_flux_y_weno_uv", "adv_flux_z_weno_uv", "adv_flux_x_weno_w",
         "adv_flux_y_weno_w", "adv_flux_z_weno_w"])
    M = _smooth_state(R, e, np.random.default_rng(seed))
    for iw, isf in pairs:
        nw, ns = R.names[iw], R.names[isf]
        out_w = int(R.args[iw, 4 + ADV_IO[nw][1]])
        out_s = int(R.args[isf, 4 + ADV_IO[ns][1]])
        assert out_w == out_s
        Mw, Ms = _run_one(R, M, iw), _run_one(R, M, isf)
        n = e.sJ * e.sK * (e.nz + 1)
        fw = Mw[out_w: out_w + n].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)
        fs = Ms[out_s: out_s + n].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)
        ks = _levels(nw, e)
        a = fw[ks, 3: 3 + e.ny + 1, 3: 3 + e.nx + 1].astype(np.float64)
        b = fs[ks, 3: 3 + e.ny + 1, 3: 3 + e.nx + 1].astype(np.float64)
        scale = np.abs(b).max()
        assert scale > 1.0, nw
        err = np.abs(a - b).max() / scale
        assert err < 2e-6, (nw, ns, err)
        # mass fluxes of both signs were exercised
        assert (b > 0).any() and (b < 0).any(), nw


def _jump_state(R, e, direction):
    """u, v, w = 10 on one side of a plane across the domain and 0 on the
    other; mass fluxes constant and positive (flow from the high side)."""
    M = R.M0.copy()
    K, J, I = np.meshgrid(np.arange(e.nz + 1), np.arange(e.sK) - 3, np.arange(e.sJ) - 3, indexing="ij")
    coord = {"x": I, "y": J, "z": K}[direction]
    mid = {"x": e.nx // 2, "y": e.ny // 2, "z": e.nz // 2}[direction]
    for name in ("u", "v", "w"):
        _field(M, e, name)[:] = np.where(coord < mid, 10.0, 0.0).astype(np.float32)
    for name in ("rru", "rrv", "rrw"):
        _field(M, e, name)[:] = 1.0
    return M


@pytest.mark.parametrize("direction", ["x", "y", "z"])
def test_momentum_weno_bounded_at_a_jump(direction):
    """Face values of the WENO momentum kernels stay within [0, 10] (to 1e-3)
    across a jump from 10 to 0, while the flx5 kernels overshoot there."""
    R, e = _program(dict(GRID, weno_momentum=1))
    M = _jump_state(R, e, direction)
    checked = 0
    for iw, isf in _pairs(R):
        nw, ns = R.names[iw], R.names[isf]
        if "_%s_" % direction not in nw:
            continue
        out = int(R.args[iw, 4 + ADV_IO[nw][1]])
        n = e.sJ * e.sK * (e.nz + 1)
        fw = _run_one(R, M, iw)[out: out + n].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)
        fs = _run_one(R, M, isf)[out: out + n].view(np.float32).reshape(e.nz + 1, e.sK, e.sJ)
        ks = slice(3, e.nz - 2) if direction == "z" else slice(1, e.nz - 1)
        a = fw[ks, 3: 3 + e.ny, 3: 3 + e.nx]
        b = fs[ks, 3: 3 + e.ny, 3: 3 + e.nx]
        assert a.max() <= 10.0 + 1e-3 and a.min() >= -1e-3, (nw, a.min(), a.max())
        assert b.max() > 10.3 or b.min() < -0.3, (ns, b.min(), b.max())
        checked += 1
    assert checked == 3  # the u and v fluxes (one kernel) and the w flux of this direction


def test_dry_diagnostics_equal_the_moist_kernel_without_moisture():
    Rd, ed = _program(dict(GRID, moist=0))
    Rm, em = _program(dict(GRID, moist=1))
    idd = Rd.names.index("update_diagnostics")
    im = Rm.names.index("update_diagnostics_moist")
    for R, e, i in ((Rd, ed, idd), (Rm, em, im)):
        assert int(R.args[i, 5]) == e.field_offset("th") and int(R.args[i, 6]) == e.field_offset("pp3d")
    Md, Mm = Rd.M0.copy(), Rm.M0.copy()
    rng = np.random.default_rng(5)
    shape = (ed.nz + 1, ed.sK, ed.sJ)
    thp = rng.normal(0, 2.0, shape).astype(np.float32)
    pp = rng.normal(0, 2e-4, shape).astype(np.float32)
    for M, e in ((Md, ed), (Mm, em)):
        _field(M, e, "th")[:] = thp
        _field(M, e, "pp3d")[:] = pp
    # the moist memory takes the dry base state; moisture fields and profiles zero
    for prof in ("th0", "pi0", "rth0"):
        od, om = ed.region_offset(prof), em.region_offset(prof)
        Mm[om: om + ed.nz] = Md[od: od + ed.nz]
    for prof in ("qv0", "qc0", "qi0"):
        om = em.region_offset(prof)
        Mm[om: om + em.nz] = 0
    for f in ("qv", "qc"):
        _field(Mm, em, f)[:] = 0.0
    Md1, Mm1 = _run_one(Rd, Md, idd), _run_one(Rm, Mm, im)
    k = slice(0, ed.nz)
    for name, ulps in (("thv", 0), ("piadv", 0), ("ppterm", 4), ("buoy", 2)):
        a = _field(Md1, ed, name)[k].astype(np.float64)
        b = _field(Mm1, em, name)[k].astype(np.float64)
        assert np.isfinite(a).all() and np.abs(a).max() > 0, name
        tol = ulps * np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
        assert (np.abs(a - b) <= tol).all(), (name, np.abs(a - b).max())


def _profile(e, M, name, n):
    o = e.region_offset(name)
    return M[o: o + n].view(np.float32).astype(np.float64)


@pytest.mark.parametrize("faces", ["tanh120", "tanh40", "uniform"])
def test_stretched_vertical_grid(faces):
    zf = {"tanh120": te_host.tanh_faces(),
          "tanh40": te_host.tanh_faces(40, 50.0, 800.0, 4000.0),
          "uniform": [500.0 * k for k in range(41)]}[faces]
    nz = len(zf) - 1
    R, e = _program(dict(nx=16, ny=12, nz=nz, zf=zf, dz=0.0, tke=1))
    M = R.M0
    Z = np.array(zf, np.float32).astype(np.float64)
    zh = ((Z[:-1] + Z[1:]) * 0.5).astype(np.float32).astype(np.float64)
    assert np.array_equal(_profile(e, M, "zf", nz + 1), Z)
    c1, c2 = _profile(e, M, "c1", nz + 1), _profile(e, M, "c2", nz + 1)
    mh, mf = _profile(e, M, "mh", nz), _profile(e, M, "mf", nz + 1)
    assert np.abs(c1 + c2 - 1.0).max() <= 6e-8
    # interior faces: c1 f(zh[k-1]) + c2 f(zh[k]) = f(zf[k]) for linear f
    for a, b in ((1.0, 2e-3), (-3.0, 5e-4)):
        lin = lambda z: a + b * z  # noqa: E731
        got = c1[1:nz] * lin(zh[:-1]) + c2[1:nz] * lin(zh[1:])
        assert np.abs(got - lin(Z[1:nz])).max() <= 1e-6 * np.abs(lin(Z)).max()
    # weights lie in [0, 1] and c2 is the face's fractional position between centres
    assert (c2[1:nz] > 0).all() and (c2[1:nz] < 1).all()
    dzn = float(Z[-1] - Z[0]) / nz
    # layer thicknesses dz_nominal / mh sum to the column depth, centre spacings to zh[-1] - zh[0]
    assert abs(np.sum(dzn / mh) - (Z[-1] - Z[0])) <= 1e-5 * Z[-1]
    assert abs(np.sum(dzn / mf[1:nz]) - (zh[-1] - zh[0])) <= 1e-5 * Z[-1]
    # Fornberg extrapolation to the surface and top faces: exact on quadratics
    # the weights travel as setBytes payloads of turb_rho_faces (parameters 6 and 7)
    A = R.args[R.names.index("turb_rho_faces")]
    cgs = R.M[int(A[4 + 6]): int(A[4 + 6]) + 3].view(np.float32).astype(np.float64)
    cgt = R.M[int(A[4 + 7]): int(A[4 + 7]) + 3].view(np.float32).astype(np.float64)
    dB, dT = Z[1] - Z[0], Z[-1] - Z[-2]
    lo, hi = zh[:3] / dB, zh[[-1, -2, -3]] / dT
    for p in range(3):
        assert abs(np.dot(cgs, lo ** p) - (Z[0] / dB) ** p) <= 2e-6 * max(1.0, (Z[0] / dB) ** p), ("surface", p)
        assert abs(np.dot(cgt, hi ** p) - (Z[-1] / dT) ** p) <= 2e-6 * (Z[-1] / dT) ** p, ("top", p)
    # rf0 at the faces: interior faces interpolate rho0, the end faces extrapolate it
    rho0, rf0 = _profile(e, R.M, "rho0", nz), _profile(e, R.M, "rf0", nz + 1)
    assert np.abs(rf0[1:nz] - (c1[1:nz] * rho0[:-1] + c2[1:nz] * rho0[1:])).max() <= 1e-6
    assert abs(rf0[0] - np.dot(cgs, rho0[:3])) <= 1e-6
    assert abs(rf0[nz] - np.dot(cgt, rho0[[-1, -2, -3]])) <= 1e-6
