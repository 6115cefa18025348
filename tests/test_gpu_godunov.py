"""GPU parity of the first-order Godunov scheme (flux_solver = GODUNOV) against the oracle:
single sweeps for both models, every boundary combination and both ghost widths, and whole
runs under each splitting; bitwise, with the per-field relative L1/Linf bar of 1e-10 stated."""
import math

import numpy as np
import pytest

import cfdart
import oracle
from util import bitwise_equal, field_errors, n_bit_diffs

pytestmark = pytest.mark.gpu
TOL = 1e-10


def interior_g(a, g):
    return a[g:-g, g:-g]


def check(gpu_state, ora_state, g, label):
    a, b = interior_g(gpu_state, g), interior_g(ora_state, g)
    errs = field_errors(a, b)
    print(f"{label}: bit-diffs={n_bit_diffs(a, b)} relL1/Linf per field={errs}", flush=True)
    for l1, li in errs:
        assert l1 <= TOL and li <= TOL, (label, errs)
    assert bitwise_equal(a, b), label


def case(model, nx, ny, bc, ng, seed):
    if model == 0:
        d, p = cfdart.configs.euler_discs_params(nx, 6, (0.05, 0.2), bc, seed, ny=ny)
        d.flux_solver, d.ng = cfdart.GODUNOV, ng
        h = cfdart.random_discs(d, p)
    else:
        d, p = cfdart.configs.allaire_bubbles_params(nx, 4, seed, ny=ny)
        d.flux_solver, d.ng = cfdart.GODUNOV, ng
        for i in range(4):
            d.bc[i] = bc[i]
        h = cfdart.shocked_bubbles(d, p)
    return d, h


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("bc", [(0, 0, 0, 0), (1, 1, 1, 1), (1, 0, 0, 1)])
@pytest.mark.parametrize("ng", [1, 2])
def test_godunov_single_sweep(gpu, model, axis, bc, ng):
    d, h = case(model, 70, 45, bc, ng, seed=7 + axis)
    cfl = 0.8 if model == 0 else 0.5
    print(f"case godunov sweep model={model} axis={axis} bc={bc} ng={ng}", flush=True)
    ora = oracle.Sim(d, h)
    dt, _ = ora.compute_dt(cfl=cfl)
    g = cfdart.Grid(d)
    g.upload(h)
    gdt, _ = g.compute_dt(0, 0.0, math.inf, cfl)
    assert gdt == dt
    ora.sweep(axis, dt)
    g.sweep(axis, dt)
    check(g.download(), ora.grid, ng, f"godunov model={model} axis={axis} bc={bc} ng={ng}")


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("splitting", [cfdart.NAIVE, cfdart.STRANG, cfdart.ALTERNATING])
def test_godunov_run(gpu, model, splitting):
    d, h = case(model, 96, 80, (1, 0, 0, 1), 1, seed=11)
    cfl = 0.8 if model == 0 else 0.5
    print(f"case godunov run model={model} splitting={splitting}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n = g.run(splitting, 40, cfl=cfl)
    ora = oracle.Sim(d, h)
    k, _ = ora.run(splitting, 40, cfl=cfl)
    assert n == k == 40 and t == ora.t
    check(g.download(), ora.grid, 1, f"godunov run model={model} splitting={splitting}")
    assert g.stats() == ora.counters()


def test_batch_refuses_godunov(gpu):
    d, h = case(0, 32, 32, (0, 0, 0, 0), 2, seed=1)
    b = cfdart.Batch(2)
    with pytest.raises(cfdart.CfdartError) as e:
        b.add(d, h, max_steps=5)
    assert e.value.code == -1
