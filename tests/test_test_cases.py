"""The settings-file driver's test-case table (cfdart_test_case, the reference's set_parameters and
set_ICs: E/sourcecode/euler_io.cpp:13-234,436-760, A/sourcecode/allaire_diffuse_io.cpp:13-310,
400-990). Every live test-case name of the reference is accepted. Each initial state is checked
against what its definition fixes independently of the code that builds it: the area of each
fluid's region (the volume fractions summed over the cells, against the closed-form area of the
circle, rectangle, sinusoidal or perturbed-circle interface), and the exact state of every cell
a half-space assigns (the oracle's primitive-to-conserved conversion of the case's post-shock
state)."""
import ctypes as C
import math

import numpy as np
import pytest

import cfdart
import oracle

EULER_CASES = ["circular_explosion", "multiple_circles_1", "multiple_circles_2", "cross", "internet", "totem_pole_1",
               "totem_pole_2", "totem_pole_3", "random_discs"]
ALLAIRE_CASES = ["ST1_x", "ST1_y", "ST2_x", "ST2_y"] + [f"TTC{k}_{a}_pure" for k in range(1, 6) for a in "xy"] + [
    "circular_explosion", "shocked_helium_bubble", "shocked_R22_bubble", "underwater_shocked_bubble",
    "underwater_explosion", "underwater_explosion_modified", "shocked_SF6", "TSTM", "RMI_SF6", "tin_air_implosion"]


def p2c(d, W):
    od = oracle.desc_from(d)
    W = np.ascontiguousarray(W, dtype=np.float64)
    U = np.zeros_like(W)
    oracle.lib().orc_p2c(C.byref(od), W.ctypes.data_as(C.POINTER(C.c_double)), U.ctypes.data_as(C.POINTER(C.c_double)))
    return U


def centres(d):
    x = d.x0 + (np.arange(d.nx) + 0.5) * d.dx
    y = d.y0 + (np.arange(d.ny) + 0.5) * d.dy
    return np.meshgrid(x, y)


@pytest.mark.parametrize("name", EULER_CASES)
def test_euler_cases_accepted(name):
    d, h, T = cfdart.test_case("euler", name, 40, 40)
    assert d.model == cfdart.EULER and T > 0 and np.all(np.isfinite(h))
    assert np.all(h[2:-2, 2:-2, 0] > 0)


@pytest.mark.parametrize("name", ALLAIRE_CASES)
def test_allaire_cases_accepted_and_physical(name):
    d, h, T = cfdart.test_case("allaire", name, 40, 32)
    assert d.model == cfdart.ALLAIRE and T > 0 and np.all(np.isfinite(h))
    U = h[2:-2, 2:-2]
    assert np.all(U[..., 0] >= 0) and np.all(U[..., 1] >= 0) and np.all(U[..., 0] + U[..., 1] > 0)
    assert np.all((U[..., 5] >= 0) & (U[..., 5] <= 1))
    od = oracle.desc_from(d)
    L = oracle.lib()
    for i in range(0, d.ny, 7):
        for j in range(0, d.nx, 5):
            c = np.ascontiguousarray(U[i, j])
            assert L.orc_is_physical(C.byref(od), c.ctypes.data_as(C.POINTER(C.c_double)))


def test_unknown_case_refused():
    with pytest.raises(cfdart.CfdartError) as e:
        cfdart.test_case("allaire", "no_such_case", 16, 16)
    assert e.value.code == -1


def area(d, h, frac):
    return float(frac(h[2:-2, 2:-2]).sum()) * d.dx * d.dy


def tin_interface_area():
    """Area of {r <= 20 + 0.4 cos 22t + 0.4 cos 17t + 0.3 cos 29t, t = atan2(y, x) - pi/2} in x >= 0:
    half the integral of r^2 over t in [-pi, 0] (midpoint rule, 2^20 points)."""
    n = 1 << 20
    t = -math.pi + (np.arange(n) + 0.5) * (math.pi / n)
    r = 20.0 + 0.4 * np.cos(22 * t) + 0.4 * np.cos(17 * t) + 0.3 * np.cos(29 * t)
    return 0.5 * float(np.sum(r * r)) * (math.pi / n)


# (case, grid, fraction whose area is fixed, closed-form area, cells masked out of the sum)
AREAS = [
    ("circular_explosion", (200, 200), lambda U: U[..., 5], math.pi * 0.4 ** 2),
    ("shocked_helium_bubble", (325, 89), lambda U: 1 - U[..., 5], math.pi * 25.0 ** 2),
    ("shocked_R22_bubble", (400, 80), lambda U: 1 - U[..., 5], math.pi * 0.025 ** 2),
    ("underwater_shocked_bubble", (240, 240), lambda U: 1 - U[..., 5], math.pi * 3.0 ** 2),
    ("shocked_SF6", (360, 160), lambda U: 1 - U[..., 5], 0.15 * 0.1),
    ("TSTM", (280, 120), lambda U: 1 - U[..., 5], 6.0 * 1.5),
    ("RMI_SF6", (400, 100), lambda U: 1 - U[..., 5], 2.9 * 0.5),  # the sine integrates to 0 over [0, 0.5]
    ("tin_air_implosion", (200, 400), lambda U: 1 - U[..., 5], None),
]


@pytest.mark.parametrize("name,grid,frac,want", AREAS, ids=[a[0] for a in AREAS])
def test_region_areas(name, grid, frac, want):
    d, h, _ = cfdart.test_case("allaire", name, *grid)
    if want is None:
        want = tin_interface_area()
    got = area(d, h, frac)
    assert abs(got - want) <= 5e-3 * want, (got, want)


def test_underwater_explosion_regions():
    """Air bubble of radius 1 at the origin under air for y >= 2.5 (cell centres): the air
    fraction sums to pi below the cut, every cell above it holds the air state exactly."""
    for name in ("underwater_explosion", "underwater_explosion_modified"):
        d, h, _ = cfdart.test_case("allaire", name, 200, 200)
        X, Y = centres(d)
        U = h[2:-2, 2:-2]
        below = Y < 2.5
        assert abs(float(U[..., 5][below].sum()) * d.dx * d.dy - math.pi) <= 5e-3 * math.pi
        air = p2c(d, [1.0, 0.0, 0.0, 0.0, 1.0e5, 1.0])
        assert np.array_equal(U[~below], np.broadcast_to(air, U[~below].shape))
    assert list(d.bc) == [0, 0, 1, 1]


@pytest.mark.parametrize("name,W,a,b,c", [
    ("shocked_helium_bubble", [1.3764, 0.0, -0.394, 0.0, 1.5698, 1.0], -1.0, 0.0, -225.0),
    ("shocked_R22_bubble", [1.686, 0.0, -113.5, 0.0, 1.59e5, 1.0], -1.0, 0.0, -0.275),
    ("underwater_shocked_bubble", [1.31, 0.0, 67.32, 0.0, 19000.0, 1.0], 1.0, 0.0, 2.4),
])
def test_post_shock_half_spaces(name, W, a, b, c):
    d, h, _ = cfdart.test_case("allaire", name, 160, 120)
    X, Y = centres(d)
    U = h[2:-2, 2:-2]
    inside = X * a + Y * b <= c
    assert inside.any() and (~inside).any()
    post = p2c(d, W)
    assert np.array_equal(U[inside], np.broadcast_to(post, U[inside].shape))
    assert not np.any(np.all(U[~inside] == post, axis=-1))


def test_case_parameters():
    """Materials, domains, end times and walls of the new cases (allaire_diffuse_io.cpp:133-310)."""
    want = {
        "circular_explosion": (1.4, 0.0, 1.4, 0.0, 0.0, 0.0, 2.0, 2.0, 0.25, [0, 0, 0, 0]),
        "underwater_shocked_bubble": (7.15, 3309.0, 1.4, 0.0, 0.0, 0.0, 12.0, 12.0, 0.05, [0, 0, 0, 0]),
        "underwater_explosion": (1.4, 0.0, 7.15, 3.309e8, -5.0, -5.0, 10.0, 10.0, 0.005, [0, 0, 0, 0]),
        "underwater_explosion_modified": (1.4, 0.0, 7.15, 3.309e8, -5.0, -5.0, 10.0, 10.0, 0.010, [0, 0, 1, 1]),
        "shocked_helium_bubble": (1.4, 0.0, 1.667, 0.0, 0.0, 0.0, 325.0, 89.0, 280.0, [0, 0, 1, 1]),
        "shocked_R22_bubble": (1.4, 0.0, 1.249, 0.0, 0.0, 0.0, 0.445, 0.089, 0.001080, [0, 0, 1, 1]),
        "shocked_SF6": (1.4, 0.0, 1.076, 0.0, 0.0, 0.0, 0.45, 0.2, 0.25, [0, 1, 1, 1]),
        "RMI_SF6": (1.4, 0.0, 1.093, 0.0, 0.0, 0.0, 4.0, 0.5, 10.0, [0, 0, 1, 1]),
        "tin_air_implosion": (3.27, 149500.0, 1.4, 0.0, 0.0, -25.0, 25.0, 50.0, 0.08, [1, 0, 0, 0]),
        "TSTM": (1.5, 0.0, 1.4, 0.0, 0.0, 0.0, 7.0, 3.0, 8.0, [0, 0, 0, 0]),
    }
    for name, (g1, p1, g2, p2, x0, y0, lx, ly, T, bc) in want.items():
        d, _, t_end = cfdart.test_case("allaire", name, 50, 40)
        assert (d.gamma1, d.pinf1, d.gamma2, d.pinf2, d.x0, d.y0) == (g1, p1, g2, p2, x0, y0), name
        assert d.dx == lx / 50 and d.dy == ly / 40 and t_end == T and list(d.bc) == bc, name
