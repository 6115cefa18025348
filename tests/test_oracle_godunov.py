"""Pins of the oracle's first-order Godunov scheme (E/headers/flux_solver_godunov.hpp:18-41,
A/headers/flux_solver_godunov.hpp:17-41, upwind z update A/headers/zupdate.hpp:40-55):

* each face flux is the HLLC flux of the two adjacent cells' stored states (the stencil of a
  first-order scheme), checked face by face against the oracle's own HLLC on a random line;
* the upwind z update, written out from the formula, on states where it differs from the
  second-order update's rounding path;
* first-order convergence to Toro's exact Riemann solution (Sod, water/air), with errors
  above the MUSCL-Hancock scheme's on the same grid;
* free-stream preservation, conservation in a closed box, and x/y transposition invariance.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from exact_riemann import Side, solve
from test_oracle_pins import dp, p2c

L = oracle.lib()
L.orc_zupdate_upwind.restype = C.c_double
L.orc_zupdate_upwind.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double]


def gdesc(**kw):
    kw.setdefault("ng", 1)
    return oracle.make_desc(flux_solver=1, **kw)


def hllc(d, UL, UR):
    n = len(UL)
    F = np.zeros(n)
    us = C.c_double()
    L.orc_hllc(C.byref(d), dp(UL), dp(UR), F.ctypes.data_as(C.POINTER(C.c_double)), C.byref(us))
    return F, us.value


@pytest.mark.parametrize("model", [0, 1])
def test_godunov_face_is_hllc_of_adjacent_cells(model):
    rng = np.random.default_rng(5 + model)
    nx = 12
    if model == 0:
        d = gdesc(model=0, nx=nx, ny=1, dx=0.1, dy=0.1, gamma1=1.4)
        W = np.column_stack([rng.uniform(0.5, 2, nx + 2), rng.uniform(-1, 1, nx + 2), rng.uniform(-1, 1, nx + 2),
                             rng.uniform(0.5, 3, nx + 2)])
    else:
        d = gdesc(model=1, nx=nx, ny=1, dx=0.1, dy=0.1, gamma1=4.4, pinf1=6e8, gamma2=1.4, pinf2=0.0)
        z = rng.uniform(0.1, 0.9, nx + 2)
        W = np.column_stack([z * 1000, (1 - z) * 1.0, rng.uniform(-5, 5, nx + 2), rng.uniform(-5, 5, nx + 2),
                             rng.uniform(1e5, 1e6, nx + 2), z])
    nc = W.shape[1]
    grid = np.zeros((3, nx + 2, nc))
    for i in range(3):
        for j in range(nx + 2):
            grid[i, j] = p2c(d, W[j])
    s = oracle.Sim(d, grid)
    row = s.fill_ghosts()[1].copy()  # the sweep's pre_sweep ghost fill (transmissive)
    dt = 1e-4 if model else 1e-3
    s.sweep(0, dt)
    r = dt / d.dx
    for j in range(1, nx + 1):
        FL, usL = hllc(d, row[j - 1], row[j])
        FR, usR = hllc(d, row[j], row[j + 1])
        expect = row[j] + r * (FL - FR)
        if model == 1:
            zL, zc, zR = row[j - 1, 5], row[j, 5], row[j + 1, 5]
            newz = zc - (dt / d.dx) * ((zR - zc) * min(0.0, usR) + (zc - zL) * max(0.0, usL))
            expect[5] = max(min(1.0, newz), 0.0)
            np.testing.assert_array_equal(s.grid[1, j, :5], expect[:5])
            assert s.grid[1, j, 5] == expect[5]
        else:
            np.testing.assert_array_equal(s.grid[1, j], expect)


def test_upwind_z_update_cases():
    d = gdesc(model=1, nx=4, ny=4, dx=0.1, dy=0.1, gamma1=4.4, pinf1=6e8, gamma2=1.4, z_lo=1e-6, z_hi=1 - 1e-6)
    f = lambda zL, z, zR, uL, uR, dt=0.01: L.orc_zupdate_upwind(C.byref(d), 0.1, dt, zL, z, zR, uL, uR)
    # flow to the right: only the left difference enters
    assert f(0.2, 0.5, 0.9, 2.0, 3.0) == 0.5 - (0.01 / 0.1) * (0.5 - 0.2) * 2.0
    # flow to the left: only the right difference
    assert f(0.2, 0.5, 0.9, -2.0, -3.0) == 0.5 - (0.01 / 0.1) * ((0.9 - 0.5) * -3.0)
    # diverging faces: neither; converging: both
    assert f(0.2, 0.5, 0.9, -1.0, 1.0) == 0.5
    assert f(0.2, 0.5, 0.9, 1.0, -1.0) == 0.5 - 0.1 * ((0.9 - 0.5) * -1.0 + (0.5 - 0.2) * 1.0)
    # clamp to the configured bounds
    assert f(0.0, 1e-6, 0.0, 0.0, -50.0) == 1e-6
    assert f(1.0, 1 - 1e-6, 1.0, 50.0, 0.0) == 1 - 1e-6


def l1_density_error(n, case, god):
    if case == "sod":
        d = oracle.make_desc(model=0, nx=n, ny=2, dx=1.0 / n, dy=1.0 / n, gamma1=1.4, flux_solver=int(god),
                             ng=1 if god else 2)
        WL, WR, T = [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1], 0.2
        Ls, Rs = Side(1.0, 0.0, 1.0, 1.4, 0.0), Side(0.125, 0.0, 0.1, 1.4, 0.0)
        rho_of = lambda U: U[..., 0]
    else:
        d = oracle.make_desc(model=1, nx=n, ny=2, dx=1.0 / n, dy=1.0 / n, gamma1=4.4, pinf1=6e8, gamma2=1.4,
                             z_lo=1e-8, z_hi=1 - 1e-8, flux_solver=int(god), ng=1 if god else 2)
        z = 1 - 1e-8
        WL = [1000.0 * z, 50.0 * (1 - z), 0.0, 0.0, 1e9, z]
        WR = [1000.0 * (1 - z), 50.0 * z, 0.0, 0.0, 1e5, 1 - z]
        T = 1e-4
        Ls, Rs = Side(1000.0, 0.0, 1e9, 4.4, 6e8), Side(50.0, 0.0, 1e5, 1.4, 0.0)
        rho_of = lambda U: U[..., 0] + U[..., 1]
    g = d.ng
    UL, UR = p2c(d, WL), p2c(d, WR)
    grid = np.zeros((2 + 2 * g, n + 2 * g, len(WL)))
    for j in range(n + 2 * g):
        grid[:, j] = UL if (j - g + 0.5) / n <= 0.5 else UR
    s = oracle.Sim(d, grid)
    s.run(0, 100000, T=T, cfl=0.5)
    xs = [(j + 0.5) / n for j in range(n)]
    ex = np.array([r[0] for r in solve(Ls, Rs, 0.5, T, xs)])
    num = rho_of(s.grid[g, g:-g])
    return np.sum(np.abs(num - ex)) / n / np.mean(ex)


@pytest.mark.parametrize("case", ["sod", "water_air"])
def test_godunov_first_order_convergence(case):
    errs = [l1_density_error(n, case, True) for n in (50, 100, 200, 400)]
    muscl = l1_density_error(400, case, False)
    print(case, errs, muscl)
    assert all(b < a for a, b in zip(errs, errs[1:]))
    rate = math.log(errs[0] / errs[-1]) / math.log(8)
    assert 0.4 < rate < 1.1  # first order (about 1/2 at discontinuities, 1 in smooth parts)
    assert muscl < errs[-1]


def test_godunov_free_stream_and_closed_box():
    for model, W in ((0, [1.3, 0.4, -0.25, 2.0]), (1, [300.0, 0.7, 15.0, -3.0, 2e6, 0.3])):
        kw = dict(gamma1=1.4) if model == 0 else dict(gamma1=4.4, pinf1=6e8, gamma2=1.4)
        d = gdesc(model=model, nx=12, ny=10, dx=0.1, dy=0.1, **kw)
        U = p2c(d, W)
        grid = np.zeros((12, 14, len(W)))
        grid[:] = U
        s = oracle.Sim(d, grid)
        s.run(2, 6, cfl=0.8)
        assert np.array_equal(s.grid[1:-1, 1:-1], grid[1:-1, 1:-1])
    # closed box: mass and energy to round-off
    d = gdesc(model=0, nx=30, ny=24, dx=1 / 30, dy=1 / 30, gamma1=1.4, bc=(1, 1, 1, 1))
    rng = np.random.default_rng(2)
    grid = np.zeros((26, 32, 4))
    for i in range(26):
        for j in range(32):
            grid[i, j] = p2c(d, [rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 3)])
    tot0 = [math.fsum(grid[1:-1, 1:-1, k].ravel()) for k in (0, 3)]
    s = oracle.Sim(d, grid)
    s.run(2, 40, cfl=0.8)
    assert s.counters()[2] == 0
    tot = [math.fsum(s.grid[1:-1, 1:-1, k].ravel()) for k in (0, 3)]
    for a, b in zip(tot0, tot):
        assert abs(a - b) <= 1e-13 * abs(a)


@pytest.mark.parametrize("model", [0, 1])
def test_godunov_x_y_transposition_bitwise(model):
    n, m = 40, 3
    if model == 0:
        kw = dict(model=0, gamma1=1.4)
        WL, WR = [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]
    else:
        kw = dict(model=1, gamma1=4.4, pinf1=6e8, gamma2=1.4, z_lo=1e-6, z_hi=1 - 1e-6)
        WL, WR = [1000 * (1 - 1e-6), 1e-6, 0.0, 0.0, 1e9, 1 - 1e-6], [1e-3, 50.0, 0.0, 0.0, 1e5, 1e-6]
    dxd = gdesc(nx=n, ny=m, dx=1.0 / n, dy=1.0 / n, **kw)
    dyd = gdesc(nx=m, ny=n, dx=1.0 / n, dy=1.0 / n, **kw)
    nc = len(WL)
    gx = np.zeros((m + 2, n + 2, nc))
    gy = np.zeros((n + 2, m + 2, nc))
    mi = 2 if model else 1
    WLy, WRy = list(WL), list(WR)
    WLy[mi], WLy[mi + 1] = WL[mi + 1], WL[mi]
    WRy[mi], WRy[mi + 1] = WR[mi + 1], WR[mi]
    for j in range(n + 2):
        left = (j - 1 + 0.5) / n <= 0.5
        gx[:, j] = p2c(dxd, WL if left else WR)
        gy[j, :] = p2c(dyd, WLy if left else WRy)
    sx, sy = oracle.Sim(dxd, gx), oracle.Sim(dyd, gy)
    sx.run(0, 15, cfl=0.5)
    sy.run(0, 15, cfl=0.5)
    ax = sx.grid[1:-1, 1:-1]
    ay = np.transpose(sy.grid[1:-1, 1:-1], (1, 0, 2)).copy()
    ay[..., [mi, mi + 1]] = ay[..., [mi + 1, mi]]
    assert np.array_equal(ax, ay)
