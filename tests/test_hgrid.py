"""Stretched horizontal grids (CPU): closed forms of CM1's stretching.

The encoder's hmetric region carries each horizontal axis (Grid.metricValues,
web/core/grid.js:200-214): map factors at centres (uh) and faces (uf), face
positions and, for x, the centres.  The recorded supercell_stretched buffer
pins one configuration word for word (test_encoder.py); these tests pin what
the stretching formula (web/core/grid.js:102-143, CM1 param.F stretch_x)
fixes for any configuration:

* the central zone has `inner` spacing; each stretched zone of n3 cells has
  spacings growing linearly from `inner` (next to the centre) to `outer`,
  so it is n3 * (inner + outer) / 2 long and the domain is tot_x_len long;
* stretch 1 is mirror-symmetric about the centre; stretch 2 keeps the west
  side at `inner` and starts at -tot_x_len / 2;
* halo cells continue at `outer` (`inner` on stretch 2's west side);
* uh = dx / (face spacing), uf = dx / (centre spacing), centres halfway;
* the automatic acoustic substep count follows the finest spacing.
"""
import numpy as np
import pytest

import te_host

# x: stretch 1, 14 + 20 + 14 cells (inner 500 over 10 km, outer 2000, 45 km);
# y: stretch 2, 24 + 16 cells (inner 500 over 12 km, outer 2000, 32 km)
GRID = dict(nx=48, ny=40, nz=12, dx=1000.0, dy=800.0, stretch_x=1, dx_inner=500.0, dx_outer=2000.0,
            nos_x_len=10000.0, tot_x_len=45000.0, stretch_y=2, dy_inner=500.0, dy_outer=2000.0,
            nos_y_len=12000.0, tot_y_len=32000.0)


def axes(**kw):
    cfg = te_host.config(**kw)
    r = te_host.Run(cfg, encode_only=True)
    sJ, sK = cfg.nx + 7, cfg.ny + 7
    m = r.read(r.region_offset("hmetric"), 9 * sJ + 3 * sK).view(np.float32).astype(np.float64)
    base = 3 * sJ + 3 * sK
    x = dict(uh=m[0:sJ], uf=m[sJ:2 * sJ], xf=m[2 * sJ + 2 * sK:3 * sJ + 2 * sK], xh=m[base + 5 * sJ:base + 6 * sJ])
    y = dict(uh=m[2 * sJ:2 * sJ + sK], uf=m[2 * sJ + sK:2 * sJ + 2 * sK], xf=m[3 * sJ + 2 * sK:3 * sJ + 3 * sK])
    names, _ = r.program()
    r.close()
    return cfg, x, y, names


def ramp(inner, outer, n3):
    return inner + (outer - inner) * np.arange(n3) / (n3 - 1)


def close(a, b, scale):
    # face positions are binary32 roundings of values up to ~scale
    return np.allclose(a, b, rtol=0, atol=scale * 2.0 ** -22)


def test_both_sides_stretching():
    cfg, x, _, _ = axes(**GRID)
    n, H = cfg.nx, 3
    f = x["xf"]
    d = np.diff(f)[H:H + n]
    want = np.concatenate([ramp(500.0, 2000.0, 14)[::-1], np.full(20, 500.0), ramp(500.0, 2000.0, 14)])
    assert close(d, want, 45000.0), np.abs(d - want).max()
    assert close(f[H + n] - f[H], 45000.0, 45000.0)
    assert close(f[H:H + n + 1], -f[H:H + n + 1][::-1], 45000.0)  # mirror symmetry
    assert close(np.diff(f)[:H], 2000.0, 45000.0) and close(np.diff(f)[H + n:], 2000.0, 45000.0)


def test_one_side_stretching():
    cfg, _, y, _ = axes(**GRID)
    n, H = cfg.ny, 3
    f = y["xf"]
    d = np.diff(f)[H:H + n]
    want = np.concatenate([np.full(24, 500.0), ramp(500.0, 2000.0, 16)])
    assert close(d, want, 32000.0), np.abs(d - want).max()
    assert close(f[H], -16000.0, 32000.0) and close(f[H + n], 16000.0, 32000.0)
    assert close(np.diff(f)[:H], 500.0, 32000.0) and close(np.diff(f)[H + n:], 2000.0, 32000.0)


@pytest.mark.parametrize("axis", ["x", "y"])
def test_map_factors(axis):
    cfg, x, y, _ = axes(**GRID)
    a, dx = (x, cfg.dx) if axis == "x" else (y, cfg.dy)
    f = a["xf"]
    # spacings from binary32 faces: relative error up to (length 2^-23) / (500 m)
    assert np.allclose(a["uh"][:-1], dx / np.diff(f), rtol=3e-5, atol=0)
    assert a["uh"][-1] == a["uh"][-2]
    c = 0.5 * (f[1:] + f[:-1])
    if axis == "x":
        assert close(a["xh"][:-1], c, 45000.0)
    assert np.allclose(a["uf"][1:-1], dx / np.diff(c), rtol=3e-5, atol=0)
    assert a["uf"][0] == 1.0 and a["uf"][-1] == 1.0  # periodic: outermost faces keep 1


def test_open_sides_copy_the_edge_face_factor():
    cfg, x, y, _ = axes(**dict(GRID, open_x=1, open_y=1))
    for a, n in ((x, cfg.nx), (y, cfg.ny)):
        assert np.all(a["uf"][:3] == a["uf"][3])
        assert np.all(a["uf"][n + 4:] == a["uf"][n + 3])


def test_uniform_axis_unchanged():
    cfg, x, y, _ = axes(nx=20, ny=16, nz=8, dx=1000.0, dy=500.0)
    assert np.all(x["uh"] == 1.0) and np.all(x["uf"] == 1.0) and np.all(y["uh"] == 1.0)
    assert np.array_equal(x["xf"], (np.arange(-3, 24) * 1000.0 - 10000.0).astype(np.float32))
    assert np.array_equal(y["xf"], (np.arange(-3, 20) * 500.0 - 4000.0).astype(np.float32))
    assert np.array_equal(x["xh"], ((np.arange(-3, 24) + 0.5) * 1000.0 - 10000.0).astype(np.float32))


def test_substeps_follow_the_finest_spacing():
    """AcousticSchedule (timestep.js:54-85): dtsm = 0.6 dmin / 360, n = round(dt / dtsm)
    made even and at least 4; stages n1 + n/2 + n.  dt = 6 s: dmin 1000 m -> n = 4
    (2 + 2 + 4 substeps); dmin 500 m -> dtsm 0.833 s, n = 8 (3 + 4 + 8)."""
    _, _, _, names = axes(nx=48, ny=40, nz=12, dx=1000.0, dy=1000.0, nsound=0, fuse=0)
    assert names.count("acoustic_uv") == 8
    _, _, _, names = axes(**dict(GRID, nsound=0, fuse=0))
    assert names.count("acoustic_uv") == 15


def test_zone_counts_must_add_up():
    with pytest.raises(te_host.TeError, match="stretch_x"):
        axes(**dict(GRID, nx=47))
    with pytest.raises(te_host.TeError, match="stretch_y"):
        axes(**dict(GRID, tot_y_len=40000.0))
    with pytest.raises(te_host.TeError, match="stretch_x"):
        axes(**dict(GRID, stretch_x=3))
