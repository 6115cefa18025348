"""GPU frame products: the device schlieren image against its definition evaluated on the host
from the downloaded state (bitwise: same IEEE operations), and the asynchronous state frame
against the downloaded state."""
import ctypes as C

import numpy as np
import pytest

import cfdart

pytestmark = pytest.mark.gpu


def host_schlieren(d, state):
    g = d.ng
    rho = state[..., 0] + (state[..., 1] if d.model == 1 else 0.0)
    rho_L, rho_R = rho[g:-g, g - 1:-g - 1], rho[g:-g, g + 1:(-g + 1) or None]
    rho_B, rho_T = rho[g - 1:-g - 1, g:-g], rho[g + 1:(-g + 1) or None, g:-g]
    gradx = (rho_R - rho_L) / (2.0 * d.dx)
    grady = (rho_T - rho_B) / (2.0 * d.dy)
    v = np.sqrt(gradx * gradx + grady * grady)
    return v / v.max(), v.max()


@pytest.mark.parametrize("model,bc", [(0, (0, 0, 0, 0)), (0, (1, 1, 0, 1)), (1, (0, 0, 0, 0)), (1, (1, 0, 1, 0))])
def test_schlieren_matches_definition(gpu, model, bc):
    if model == 0:
        d, h = cfdart.configs.euler_discs(150, 8, (0.03, 0.1), bc, seed=3, ny=110)
        cfl = 0.8
    else:
        d, h = cfdart.configs.allaire_bubbles(150, 6, seed=3, ny=110)
        for i in range(4):
            d.bc[i] = bc[i]
        cfl = 0.5
    print(f"case schlieren model={model} bc={bc}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 7, cfl=cfl)
    img, m = g.schlieren()
    ref, mref = host_schlieren(d, g.download())
    assert m == mref
    assert np.array_equal(img, ref)


def test_state_frame_matches_download(gpu):
    d, h = cfdart.configs.allaire_bubbles(90, 4, seed=8, ny=70)
    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 5, cfl=0.5)
    n = d.nx * d.ny * 6
    p = cfdart.lib().cfdart_host_alloc(8 * n)
    assert p
    try:
        assert cfdart.lib().cfdart_frame(g.h, C.c_void_p(p)) == 0
        g.sync()
        frame = np.ctypeslib.as_array((C.c_double * n).from_address(p)).reshape(6, d.ny, d.nx).copy()
    finally:
        cfdart.lib().cfdart_host_free(C.c_void_p(p))
    s = g.download()[2:-2, 2:-2]
    assert np.array_equal(frame, np.moveaxis(s, 2, 0))


@pytest.mark.parametrize("model", [0, 1])
def test_mass_history_matches_compute_dt(gpu, model):
    """The device mass history of a run equals, step by step, the masses compute_dt records (the
    library's own cfdart_compute_dt bitwise: same partials and order) and the oracle's row-order
    sums to round-off; with reflective walls all round the masses stay constant to round-off."""
    import math

    import oracle
    if model == 0:
        d, h = cfdart.configs.euler_discs(120, 6, (0.05, 0.15), (1, 1, 1, 1), seed=4, ny=90)
        cfl = 0.8
    else:
        d, h = cfdart.configs.allaire_bubbles(120, 5, seed=4, ny=90)
        cfl = 0.5
    print(f"case mass history model={model}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n, hist = g.run_mass(cfdart.ALTERNATING, 25, cfl=cfl)
    assert n == 25 and hist.shape == (25, 3)
    ora = oracle.Sim(d, h)
    step_g = cfdart.Grid(d)
    step_g.upload(h)
    for k in range(25):
        dt, (m1, m2) = ora.compute_dt(cfl=cfl)
        gdt, gm = step_g.compute_dt(k, ora.t, math.inf, cfl)
        assert hist[k, 0] == ora.t
        assert hist[k, 1] == gm[0] and (model == 0 or hist[k, 2] == gm[1])
        assert abs(hist[k, 1] - m1) <= 1e-13 * abs(m1)
        if model == 1:
            assert abs(hist[k, 2] - m2) <= 1e-13 * abs(m2)
        ora.advance(cfdart.ALTERNATING, dt)
        step_g.step(cfdart.ALTERNATING, k, gdt)
        ora.t += dt
        ora.n += 1
    if model == 0:  # closed box: mass conserved to round-off
        assert np.max(np.abs(hist[:, 1] - hist[0, 1])) <= 1e-13 * hist[0, 1]
