"""The caller's sounding (buildRun(configuration, sounding), isnd = 7 / 17), CPU.

te_config.input_sounding takes a CM1 input_sounding file; the base state is
Sounding + Sounding.interpolated + BaseState.initProfiles
(web/core/base.js:659-757, :148-178).  No recording of the reference holds
the profiles of a file sounding, so the path is pinned by what fixes them:

* linear theta, u, v in the file -> linear profiles on the grid, the surface
  winds extrapolated from the first two rows;
* a dry isentropic file -> pi(z) = pi_sfc - g z / (cp theta) exactly (the
  trapezoid rule is exact for constant theta_v);
* grid levels that coincide with the file's rows -> the file's theta and qv
  (qv through relative humidity and back: a few ulp);
* rows drier than 1e-12 get 0.1 % of saturation; isnd = 17 has no winds;
* malformed files and soundings below the grid top are refused.
"""
import os

import numpy as np
import pytest

import te_host

HERE = os.path.dirname(os.path.abspath(__file__))
HURRICANE = os.path.join(HERE, "golden", "soundings", "hurricane_3d.input_sounding")
G, CP, RD, RV, P00 = 9.81, 1005.7, 287.04, 461.5, 1.0e5


def text(psfc, thsfc, qvsfc, rows, winds=True):
    s = "%12.4f %12.6f %12.6f\n" % (psfc, thsfc, qvsfc)
    for r in rows:
        s += ("%12.4f %12.6f %12.6f %10.4f %10.4f\n" % r) if winds else ("%12.4f %12.6f %12.6f\n" % r[:3])
    return s


def profiles(snd, winds=1, **kw):
    base = dict(nx=8, ny=8, nz=20, dz=500.0, moist=1, input_sounding=snd, sounding_winds=winds)
    base.update(kw)
    r = te_host.Run(te_host.config(**base), encode_only=True)
    nz = r.nz
    out = {n: r.read(r.region_offset(n), nz).view(np.float32).astype(np.float64)
           for n in ("th0", "qv0", "pi0", "u0", "v0", "rho0")}
    out["zh"] = (np.arange(nz) + 0.5) * 500.0
    r.close()
    return out


def test_linear_profiles_stay_linear():
    rows = [(z, 300.0 + 0.004 * z, 10.0, 2.0 + 0.001 * z, -1.0 + 0.0005 * z) for z in np.arange(500.0, 12001.0, 500.0)]
    p = profiles(text(1000.0, 300.0, 12.0, rows), umove=3.0, vmove=-2.0)
    zh = p["zh"]
    assert np.allclose(p["th0"], 300.0 + 0.004 * zh, rtol=0, atol=1e-4)
    # zh = 250 lies between the surface level (winds extrapolated: u(0) = 2) and z = 500
    assert np.allclose(p["u0"], 2.0 + 0.001 * zh - 3.0, rtol=0, atol=1e-5)
    assert np.allclose(p["v0"], -1.0 + 0.0005 * zh + 2.0, rtol=0, atol=1e-5)


def test_dry_isentropic_pi_is_linear():
    rows = [(z, 300.0, 0.0, 0.0, 0.0) for z in np.arange(400.0, 15001.0, 400.0)]
    p = profiles(text(1000.0, 300.0, 0.0, rows), moist=0)
    assert np.all(p["th0"] == np.float32(300.0)) and np.all(p["qv0"] == 0.0)
    want = 1.0 - np.float32(G) * p["zh"] / (np.float32(CP) * 300.0)
    assert np.allclose(p["pi0"], want, rtol=0, atol=2e-6), np.abs(p["pi0"] - want).max()
    prs = P00 * p["pi0"] ** (CP / RD)
    assert np.allclose(p["rho0"], prs / (RD * 300.0 * p["pi0"]), rtol=2e-6)


def test_rows_on_grid_levels_come_back():
    zh = (np.arange(20) + 0.5) * 500.0
    th = 300.0 + 0.003 * zh + 2.0 * np.sin(zh / 3000.0)
    qv = 14.0 * np.exp(-zh / 2500.0)
    rows = [(z, t, q, 0.0, 0.0) for z, t, q in zip(zh, th, qv)]
    rows.append((10500.0, th[-1] + 1.5, qv[-1] * 0.5, 0.0, 0.0))
    snd = text(1000.0, 300.0, 14.5, rows)
    p = profiles(snd)
    filed = np.array([[float(t) for t in l.split()] for l in snd.splitlines()[1:21]])  # as printed
    assert np.allclose(p["th0"], np.float32(filed[:, 1]), rtol=1e-7, atol=0)
    q32 = np.float32(np.float32(filed[:, 2]) / np.float32(1000.0))
    # qv -> relative humidity at the row's (p, T) -> qv at the level's (p, T),
    # re-interpolated to within an ulp of T: es'(T)/es 1 ulp ~ 2e-6
    assert np.allclose(p["qv0"], q32, rtol=3e-6, atol=0), np.abs(p["qv0"] / q32 - 1).max()


def test_dry_rows_get_a_thousandth_of_saturation_and_isnd17_has_no_wind():
    zh = (np.arange(20) + 0.5) * 500.0
    rows = [(z, 300.0 + 0.004 * z, 0.0, 5.0, 5.0) for z in zh]
    rows.append((11000.0, 344.0, 0.0, 5.0, 5.0))
    p = profiles(text(1000.0, 300.0, 0.0, rows, winds=False), winds=0, umove=1.5)
    assert np.all(p["u0"] == np.float32(-1.5)) and np.all(p["v0"] == 0.0)
    # rh = 0.001 everywhere: qv0 = 0.001 qvs(p, T); qvs from Bolton's es over water
    T = p["th0"] * p["pi0"]
    prs = P00 * p["pi0"] ** (CP / RD)
    es = 611.2 * np.exp(17.67 * (T - 273.15) / (T - 29.65))
    qvs = (RD / RV) * es / (prs - es)
    assert np.allclose(p["qv0"], 0.001 * qvs, rtol=2e-5)


def test_reference_sounding_file():
    """The reference's own hurricane_3d sounding: theta on the grid is the
    file's theta interpolated in z; pi falls, qv stays positive."""
    rows = [list(map(float, l.split())) for l in open(HURRICANE).read().strip().splitlines()]
    z = np.array([0.0] + [r[0] for r in rows[1:]])
    th = np.array([rows[0][1]] + [r[1] for r in rows[1:]])
    p = profiles(open(HURRICANE).read(), nz=40, dz=400.0)
    zh = (np.arange(40) + 0.5) * 400.0
    assert np.allclose(p["th0"], np.interp(zh, z, th), rtol=0, atol=2e-4)
    assert np.all(np.diff(p["pi0"]) < 0) and np.all(p["qv0"] > 0)


@pytest.mark.parametrize("bad,match", [
    ("", "empty"),
    ("1000 300\n500 301 10 0 0\n1000 302 9 0 0\n", "header"),
    ("1000 300 14\n500 301 10 0\n1000 302 9 0 0\n", "line 2"),
    ("1000 300 14\n500 301 10 0 0\n", "at least 2"),
    ("1000 300 14\n500 301 10 0 0\n400 302 9 0 0\n", "increase"),
    ("1000 300 14\n500 301 10 0 0\n1000 302 9 0 0\n", "below the top"),
])
def test_bad_soundings_are_refused(bad, match):
    with pytest.raises(te_host.TeError, match=match):
        profiles(bad)


def test_context_keeps_its_copy():
    """te_create / te_encode copy the text: the caller may free it."""
    rows = [(z, 300.0 + 0.004 * z, 10.0, 1.0, 0.0) for z in np.arange(500.0, 12001.0, 500.0)]
    snd = text(1000.0, 300.0, 12.0, rows)
    cfg = te_host.config(nx=8, ny=8, nz=20, dz=500.0, input_sounding=snd, sounding_winds=1)
    r = te_host.Run(cfg, encode_only=True)
    cfg._keep.clear()
    cfg.input_sounding = b"1000 250 1\n100 250 1 0 0\n20000 250 1 0 0\n"  # the caller reuses its memory
    r.init_state()  # the t = 0 state from the context's copy of the sounding
    fresh = te_host.Run(te_host.config(nx=8, ny=8, nz=20, dz=500.0, input_sounding=snd, sounding_winds=1),
                        encode_only=True)
    fresh.init_state()
    assert np.array_equal(r.read(0, r.mem_words), fresh.read(0, fresh.mem_words))
