"""Pins of the oracle's functions that the golden tape does not fully fix
(CPU): closed forms, textbook routines and brute force on small inputs.

Each check is chosen so that a plausible slip (a dropped term, a wrong index
or sign, a transposed operand) fails it; none re-types the function under
test.
"""
import math
import struct

import numpy as np
import pytest

import oracle_lib as ol


def f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


def ulp_diff(a, b):
    ia = np.array([a], np.float32).view(np.int32)[0]
    ib = np.array([b], np.float32).view(np.int32)[0]
    return abs(int(ia) - int(ib))


# ---------------------------------------------------------------- pow / log / exp
def test_rr_pow_against_fp64_pow():
    """The double-float pow (pressure_tendency.wgsl:114-227) is correctly
    rounded where FP64 pow rounded to FP32 is: exact on >= 99.9 % of samples,
    never off by more than 1 ulp (the FP64 reference itself is within
    2^-53 relative of the true value)."""
    L = ol.lib()
    rng = np.random.default_rng(7)
    xs = rng.uniform(0.2, 1.5, 20000).astype(np.float32)
    ys = np.concatenate([np.full(10000, f32(287.04 / 718.66003), np.float32),
                         rng.uniform(0.1, 3.6, 10000).astype(np.float32)])
    exact = worst = 0
    for x, y in zip(xs, ys):
        got = L.oracle_rr_pow(float(x), float(y))
        want = f32(math.pow(float(x), float(y)))
        d = ulp_diff(got, want)
        exact += d == 0
        worst = max(worst, d)
    assert worst <= 1
    assert exact >= 0.999 * len(xs)


def test_rr_log_exp_against_fp64():
    L = ol.lib()
    rng = np.random.default_rng(11)
    for x in rng.uniform(1e-3, 1e3, 5000).astype(np.float32):
        assert ulp_diff(L.oracle_rr_log(float(x)), f32(math.log(float(x)))) <= 1
    for x in rng.uniform(-20, 20, 5000).astype(np.float32):
        assert ulp_diff(L.oracle_rr_exp(float(x)), f32(math.exp(float(x)))) <= 1
    assert L.oracle_rr_pow(1.0, 0.37) == 1.0
    assert L.oracle_rr_pow(4.0, 0.5) == 2.0


# ---------------------------------------------------------------- interpolants
def test_flx5_flx3_reproduce_polynomials():
    """flx5 is the 5th-order upwind face value: exact for polynomials up to
    degree 4 sampled at cell averages; flx3 exact up to degree 2.  The face
    lies between the 3rd and 4th of flx5's inputs (2nd and 3rd of flx3's)."""
    L = ol.lib()
    for deg in range(5):
        coeff = np.random.default_rng(deg).normal(size=deg + 1)
        P = np.polynomial.Polynomial(coeff).integ()
        cells = [(P(c + 1) - P(c)) for c in range(-3, 2)]  # cells [-3,-2] .. [1,2], face at 0
        got = L.oracle_flx5(*[float(v) for v in cells])
        want = np.polynomial.Polynomial(coeff)(0.0)
        assert abs(got - want) <= 2e-5 * (1 + abs(want)), (deg, got, want)
        if deg <= 2:
            got3 = L.oracle_flx3(*[float(v) for v in cells[1:4]])
            assert abs(got3 - want) <= 2e-5 * (1 + abs(want)), (deg, got3, want)
    # a degree-5 polynomial is not reproduced (the order is exactly 5)
    P = np.polynomial.Polynomial([0, 0, 0, 0, 0, 1.0]).integ()
    cells = [(P(c + 1) - P(c)) for c in range(-3, 2)]
    assert abs(L.oracle_flx5(*[float(v) for v in cells])) > 1e-3


def test_weno5_constant_smooth_and_jump():
    L = ol.lib()
    eps = 1e-18
    for c in (0.0, 1.0, -3.5, 300.0):
        assert L.oracle_weno5(c, c, c, c, c, eps) == pytest.approx(c, rel=1e-6, abs=1e-7)
    # smooth data: WENO5 tends to the linear 5th-order scheme (flx5)
    P = np.polynomial.Polynomial([0.3, 0.2, -0.1, 0.05]).integ()
    h = 0.05
    cells = [float((P((c + 1) * h) - P(c * h)) / h) for c in range(-3, 2)]
    w = L.oracle_weno5(*cells, eps)
    f = L.oracle_flx5(*cells)
    assert abs(w - f) < 1e-5
    # a jump: the face value stays within the data range (no new extremum)
    jump = [0.0, 0.0, 0.0, 1.0, 1.0]
    assert 0.0 <= L.oracle_weno5(*jump, eps) <= 1.0
    # the FP32 and FP64 instantiations agree to FP32 rounding
    rng = np.random.default_rng(3)
    for _ in range(2000):
        s = rng.normal(size=5)
        a = L.oracle_weno5(*[float(f32(v)) for v in s], 1e-20)
        b = L.oracle_weno5_f64(*[float(f32(v)) for v in s], 1e-20)
        assert abs(a - b) <= 2e-5 * (1 + abs(b))


def test_upstrpd_limits():
    """Flux-limited 2nd order: upwind (1st order) value at a local extremum,
    bounded by its neighbours on monotone data."""
    L = ol.lib()
    rng = np.random.default_rng(5)
    for _ in range(500):
        a, b, c = sorted(rng.normal(size=3))
        for s in ((a, b, c), (c, b, a)):
            v = L.oracle_upstrpd(float(s[0]), float(s[1]), float(s[2]), 1.0)
            assert min(s[1], s[2]) - 1e-6 <= v <= max(s[1], s[2]) + 1e-6


# ---------------------------------------------------------------- tridiagonal
def _w_solve_memory(nx, ny, nz, rng):
    sJ, sK = nx + 7, ny + 7
    fw = sJ * sK * (nz + 1)
    names = ["pk1", "pk2", "qk", "mm", "fpk", "wten", "pp3d", "w", "cA", "cB"]
    base = 64
    offs = {n: base + 16 + q * fw for q, n in enumerate(names)}
    M = np.zeros(base + 16 + len(names) * fw + sJ * sK * 4, np.uint32)
    M[0:5] = [nx, ny, nz, sJ, sK]
    M[5:8] = np.array([1e-3, 1e-3, 2e-3], np.float32).view(np.uint32)
    M[8:13] = np.array([1.0, 0.6, 0.4, 0.1, 0.0], np.float32).view(np.uint32)  # dts alph beta kdiv smeps
    F = {}
    for n in names:
        v = rng.normal(size=fw).astype(np.float32)
        if n == "mm":
            v = np.abs(v) + 0.5
        if n in ("pk1", "pk2", "qk"):
            v *= 0.3
        M[offs[n]: offs[n] + fw] = v.view(np.uint32)
    A = np.zeros((1, 64), np.uint32)
    A[0, 0:3] = [nx, ny, 1]
    A[0, 4] = 0
    A[0, 5] = 8
    for q, n in enumerate(names):
        A[0, 4 + 2 + q] = offs[n]
    return M, A, offs, sJ, sK, fw


def test_w_solve_against_dense_solve():
    """acoustic_w_solve's Thomas sweep equals numpy.linalg.solve of the same
    tridiagonal system (FP64) on small columns, with w = 0 at k = 0 and nz."""
    rng = np.random.default_rng(1)
    for nz in (3, 4, 6, 9):
        nx, ny = 3, 2
        M, A, offs, sJ, sK, fw = _w_solve_memory(nx, ny, nz, rng)
        F = {n: M[o: o + fw].view(np.float32).astype(np.float64).copy() for n, o in offs.items()}
        ol.Program(["acoustic_w_solve"], A).run(M)
        w = M[offs["w"]: offs["w"] + fw].view(np.float32)
        dts, alph, beta = 1.0, float(f32(0.6)), float(f32(0.4))
        asq = alph * alph
        for j in range(ny):
            for i in range(nx):
                idx = lambda k: (i + 3) + sJ * ((j + 3) + sK * k)
                n = nz - 1
                T = np.zeros((n, n))
                d = np.zeros(n)
                for r, k in enumerate(range(1, nz)):
                    c, cm = idx(k), idx(k - 1)
                    m = F["mm"][c]
                    aa = -asq * m * (F["pk1"][cm] + F["qk"][cm])
                    cc = asq * m * (F["pk2"][c] - F["qk"][c])
                    bb = 1 - asq * m * (-(F["pk1"][c] + F["qk"][c]) + (F["pk2"][cm] - F["qk"][cm]))
                    fw_ = dts * F["wten"][c] + beta * (-m * (F["pp3d"][c] - F["pp3d"][cm]))
                    d[r] = alph * (-m * ((F["pp3d"][c] + F["fpk"][c]) - (F["pp3d"][cm] + F["fpk"][cm]))) + F["w"][c] + fw_
                    T[r, r] = bb
                    if r > 0:
                        T[r, r - 1] = aa
                    if r < n - 1:
                        T[r, r + 1] = cc
                x = np.linalg.solve(T, d)
                got = np.array([w[idx(k)] for k in range(1, nz)], np.float64)
                scale = np.abs(x).max() + 1
                assert np.abs(got - x).max() <= 2e-5 * scale * np.linalg.cond(T), (nz, i, j)
                assert w[idx(0)] == 0 and w[idx(nz)] == 0


# ---------------------------------------------------------------- hashes
def _region_hash_py(words):
    s = 0
    for i, v in enumerate(words):
        h = (int(v) ^ ((i * 0x9E3779B1) & 0xFFFFFFFF)) & 0xFFFFFFFF
        h = (h * 0x85EBCA6B) & 0xFFFFFFFF
        h ^= h >> 13
        h = (h * 0xC2B2AE35) & 0xFFFFFFFF
        h ^= h >> 16
        s = (s + h) & 0xFFFFFFFF
    return s


def test_hash_formulas_on_hand_vectors():
    # FNV-1a over u32 words: empty input is the offset basis
    assert ol.fnv1a(np.zeros(0, np.uint32)) == 0x811C9DC5
    assert ol.fnv1a(np.array([0], np.uint32)) == (0x811C9DC5 * 0x01000193) & 0xFFFFFFFF
    # region hash of zeros: sum of mix(i * golden) only
    w = np.zeros(7, np.uint32)
    assert ol.region_hash(w, 0, 7) == _region_hash_py(w)
    w = np.arange(1000, dtype=np.uint32) * np.uint32(2654435761)
    assert ol.region_hash(w, 0, 1000) == _region_hash_py(w)
    # order matters through the index term
    assert ol.region_hash(w[::-1].copy(), 0, 1000) != ol.region_hash(w, 0, 1000)
