"""GPU parity on ragged and extreme grid shapes: the smallest grids the scheme allows (2 cells
per axis), sizes one past the x tile (60 columns) and the y block (256 columns), tall and
flat grids, for both models and both schemes; whole runs against the oracle, bitwise."""
import numpy as np
import pytest

import cfdart
import oracle
from util import bitwise_equal

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2), (3, 5), (61, 7), (257, 3), (5, 300), (121, 257)]


def state(model, nx, ny, scheme, seed):
    rng = np.random.default_rng(seed)
    if model == 0:
        d = cfdart.make_desc(cfdart.EULER, nx, ny, lx=1.0, ly=ny / nx, gamma1=1.4, bc=(1, 0, 1, 0),
                             limiter=cfdart.VANLEER, flux_solver=scheme)
        W = np.stack([rng.uniform(0.5, 2.0, (ny, nx)), rng.uniform(-0.5, 0.5, (ny, nx)),
                      rng.uniform(-0.5, 0.5, (ny, nx)), rng.uniform(0.5, 3.0, (ny, nx))], -1)
    else:
        d = cfdart.make_desc(cfdart.ALLAIRE, nx, ny, lx=1.0, ly=ny / nx, gamma1=4.4, pinf1=6e8, gamma2=1.4,
                             bc=(0, 1, 0, 1), wave_speeds=cfdart.PVRS, z_lo=1e-6, z_hi=1 - 1e-6, flux_solver=scheme)
        z = rng.uniform(0.01, 0.99, (ny, nx))
        W = np.stack([z * 1000.0, (1 - z) * 1.0, rng.uniform(-5, 5, (ny, nx)), rng.uniform(-5, 5, (ny, nx)),
                      rng.uniform(1e5, 1e6, (ny, nx)), z], -1)
    g = d.ng
    h = np.zeros((ny + 2 * g, nx + 2 * g, W.shape[-1]))
    U = np.zeros(W.shape[-1])
    for i in range(ny):
        for j in range(nx):
            cfdart.lib().cfdart_prim_to_cons(cfdart._lib.C.byref(d), cfdart._lib._dptr(np.ascontiguousarray(W[i, j])),
                                             cfdart._lib._dptr(U))
            h[i + g, j + g] = U
    return d, h


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("scheme", [cfdart.MUSCL_HANCOCK, cfdart.GODUNOV])
def test_shape_run(gpu, nx, ny, model, scheme):
    d, h = state(model, nx, ny, scheme, seed=nx * 1000 + ny)
    cfl = 0.8 if model == 0 else 0.5
    print(f"case shape {nx}x{ny} model={model} scheme={scheme}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n = g.run(cfdart.ALTERNATING, 12, cfl=cfl)
    ora = oracle.Sim(d, h)
    k, _ = ora.run(cfdart.ALTERNATING, 12, cfl=cfl)
    assert n == k == 12 and t == ora.t
    ng = d.ng
    assert bitwise_equal(g.download()[ng:-ng, ng:-ng], ora.grid[ng:-ng, ng:-ng])
    assert g.stats() == ora.counters()


def test_batch_ragged_shapes_equal_alone(gpu):
    """All the shapes above, both models, in one batch: each grid ends bitwise where it ends
    alone."""
    cases = [state(model, nx, ny, cfdart.MUSCL_HANCOCK, seed=7 * nx + ny) for nx, ny in SHAPES for model in (0, 1)]
    print(f"case batch of {len(cases)} ragged grids", flush=True)
    b = cfdart.Batch(len(cases))
    slots = [b.add(d, h, max_steps=15, cfl=0.8 if d.model == 0 else 0.5) for d, h in cases]
    b.step(16)
    b.sync()
    steps, flags = b.poll(len(cases))
    for (d, h), s in zip(cases, slots):
        assert steps[s] == 15 and flags[s] == 0
        g = cfdart.Grid(d)
        g.upload(h)
        g.run(cfdart.ALTERNATING, 15, cfl=0.8 if d.model == 0 else 0.5)
        assert g.checksum() == b.checksum(s), (d.model, d.nx, d.ny)


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("yt", [cfdart.Y_TRANSPOSED, cfdart.Y_TRANSPOSED_NARROW, 5, 33])
@pytest.mark.parametrize("xr", [1, 3, 7, 16])
def test_tile_shapes(gpu, model, yt, xr):
    """Every tile shape the autotuner may pick gives the oracle's bits (a grid that no tile
    divides evenly)."""
    d, h = state(model, 150, 97, cfdart.MUSCL_HANCOCK, seed=31 + model)
    cfl = 0.8 if model == 0 else 0.5
    print(f"case tile y={yt} x={xr} model={model}", flush=True)
    g = cfdart.Grid(d)
    g.set_tile(yt, xr)
    g.upload(h)
    t, n = g.run(cfdart.ALTERNATING, 6, cfl=cfl)
    ora = oracle.Sim(d, h)
    k, _ = ora.run(cfdart.ALTERNATING, 6, cfl=cfl)
    assert n == k == 6 and t == ora.t
    assert bitwise_equal(g.download()[2:-2, 2:-2], ora.grid[2:-2, 2:-2])
    assert g.stats() == ora.counters()
