"""GPU parity of the transposed y-sweep tiles (cfdart_set_tile with CFDART_Y_TRANSPOSED or
CFDART_Y_TRANSPOSED_NARROW): a workgroup stages 64 rows x 16 (or 8) columns through LDS and runs
lanes along y. They must give the same bits and the same fallback/repair counters as the oracle
and as the strip tile: single y sweeps over every wall combination, whole runs on shapes one past
the tile (17 or 9 columns, 61 rows) and smaller than it, and a 2048^2 Euler run compared with the
strip tile."""
import math

import pytest

import cfdart
import oracle
from test_gpu_shapes import state
from util import bitwise_equal, interior

pytestmark = pytest.mark.gpu
YT = cfdart.Y_TRANSPOSED
WIDTHS = [cfdart.Y_TRANSPOSED, cfdart.Y_TRANSPOSED_NARROW]


@pytest.mark.parametrize("bc", [(0, 0, 0, 0), (1, 1, 1, 1), (0, 1, 1, 0)])
@pytest.mark.parametrize("limiter", [cfdart.MINMOD, cfdart.VANLEER])
@pytest.mark.parametrize("yt", WIDTHS)
def test_euler_y_sweep(gpu, bc, limiter, yt):
    d, h = cfdart.configs.euler_discs(70, 6, (0.05, 0.2), bc, 17, limiter=limiter, ny=125)
    print(f"case transposed ({yt}) euler y sweep bc={bc} limiter={limiter}", flush=True)
    ora = oracle.Sim(d, h)
    dt, _ = ora.compute_dt(cfl=0.8)
    g = cfdart.Grid(d)
    g.set_tile(yt, 4)
    g.upload(h)
    ora.sweep(1, dt)
    g.sweep(1, dt)
    assert bitwise_equal(interior(g.download()), interior(ora.grid))


@pytest.mark.parametrize("ws", [cfdart.GARRICK, cfdart.PVRS])
@pytest.mark.parametrize("yt", WIDTHS)
def test_allaire_y_sweep(gpu, ws, yt):
    d, h = cfdart.configs.allaire_bubbles(50, 4, seed=9, ny=130, wave_speeds=ws)
    print(f"case transposed ({yt}) allaire y sweep ws={ws}", flush=True)
    ora = oracle.Sim(d, h)
    dt, _ = ora.compute_dt(cfl=0.5)
    g = cfdart.Grid(d)
    g.set_tile(yt, 2)
    g.upload(h)
    ora.sweep(1, dt)
    g.sweep(1, dt)
    assert bitwise_equal(interior(g.download()), interior(ora.grid))


@pytest.mark.parametrize("nx,ny", [(2, 2), (17, 61), (16, 60), (9, 121), (33, 7), (5, 181)])
@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("scheme", [cfdart.MUSCL_HANCOCK, cfdart.GODUNOV])
@pytest.mark.parametrize("yt", WIDTHS)
def test_run_shapes(gpu, nx, ny, model, scheme, yt):
    d, h = state(model, nx, ny, scheme, seed=nx * 997 + ny)
    cfl = 0.8 if model == 0 else 0.5
    print(f"case transposed ({yt}) run {nx}x{ny} model={model} scheme={scheme}", flush=True)
    g = cfdart.Grid(d)
    g.set_tile(yt, 4)
    g.upload(h)
    t, n = g.run(cfdart.STRANG, 10, cfl=cfl)
    ora = oracle.Sim(d, h)
    k, _ = ora.run(cfdart.STRANG, 10, cfl=cfl)
    assert n == k == 10 and t == ora.t
    ng = d.ng
    assert bitwise_equal(g.download()[ng:-ng, ng:-ng], ora.grid[ng:-ng, ng:-ng])
    assert g.stats() == ora.counters()


def test_euler_2048_equals_strip(gpu):
    """2048^2 Euler discs, 100 alternating steps: transposed tile == strip tile, bitwise, with the
    same counters and time."""
    d, h = cfdart.configs.euler_2048()
    out = []
    for tile in [(YT, 4), (cfdart.Y_TRANSPOSED_NARROW, 4), (24, 6)]:
        print(f"case euler_2048 100 steps tile={tile}", flush=True)
        g = cfdart.Grid(d)
        g.set_tile(*tile)
        g.upload(h)
        t, n = g.run(cfdart.ALTERNATING, 100, cfl=0.8)
        out.append((t, n, g.checksum(), g.stats()))
        g.close()
    assert out[0] == out[1] == out[2]
    assert out[0][1] == 100 and math.isfinite(out[0][0])


def test_rejects_other_negative_rows(gpu):
    d, h = cfdart.configs.euler_discs(32, 2, (0.05, 0.2), (0, 0, 0, 0), 1, ny=32)
    g = cfdart.Grid(d)
    with pytest.raises(cfdart.CfdartError):
        g.set_tile(-3, 4)
    for yt in WIDTHS:
        g.set_tile(yt, 4)
    g.close()
    # a batch runs the 16-column transposed y tile only
    b = cfdart.Batch(2)
    for y in (cfdart.Y_TRANSPOSED_NARROW, 16):
        with pytest.raises(cfdart.CfdartError):
            b.add(d, h, max_steps=2, y_strip_rows=y)
    s = b.add(d, h, max_steps=2, y_strip_rows=YT)
    with pytest.raises(cfdart.CfdartError):
        b.set_tile(s, 32, 4)
    b.set_tile(s, YT, 4)
    b.close()
