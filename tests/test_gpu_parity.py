"""GPU parity: the HIP sweeps against the CPU oracle on the same seeded inputs.

The bar is bitwise equality (both sides compute IEEE binary64 with the same operation order
and no FMA contraction); every test also states the per-field relative L1/Linf and checks
them against the 1e-10 tolerance of the benchmark definition.
"""
import math

import numpy as np
import pytest

import cfdart
import oracle
from util import bitwise_equal, field_errors, interior, n_bit_diffs

pytestmark = pytest.mark.gpu
TOL = 1e-10


def check(gpu_state, ora_state, label, bitwise=True):
    a, b = interior(gpu_state), interior(ora_state)
    errs = field_errors(a, b)
    print(f"{label}: bit-diffs={n_bit_diffs(a, b)} relL1/Linf per field={errs}", flush=True)
    for l1, li in errs:
        assert l1 <= TOL and li <= TOL, (label, errs)
    if bitwise:
        assert bitwise_equal(a, b), label


def euler_small(nx, ny, bc, limiter, seed):
    d, h = cfdart.configs.euler_discs(nx, 6, (0.05, 0.2), bc, seed, limiter=limiter, ny=ny)
    return d, h


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("bc", [(0, 0, 0, 0), (1, 1, 1, 1), (1, 0, 0, 1)])
@pytest.mark.parametrize("limiter", [cfdart.MINMOD, cfdart.VANLEER])
def test_euler_single_sweep(gpu, axis, bc, limiter):
    d, h = euler_small(70, 45, bc, limiter, seed=3 + axis)
    print(f"case euler sweep axis={axis} bc={bc} limiter={limiter}", flush=True)
    ora = oracle.Sim(d, h)
    dt, _ = ora.compute_dt(cfl=0.8)
    g = cfdart.Grid(d)
    g.upload(h)
    gdt, _ = g.compute_dt(0, 0.0, math.inf, 0.8)
    assert gdt == dt
    ora.sweep(axis, dt)
    g.sweep(axis, dt)
    check(g.download(), ora.grid, f"euler axis={axis} bc={bc} lim={limiter}")


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("ws", [cfdart.GARRICK, cfdart.PVRS])
def test_allaire_single_sweep(gpu, axis, ws):
    d, h = cfdart.configs.allaire_bubbles(64, 4, seed=5, ny=48, wave_speeds=ws)
    print(f"case allaire sweep axis={axis} ws={ws}", flush=True)
    ora = oracle.Sim(d, h)
    dt, _ = ora.compute_dt(cfl=0.5)
    g = cfdart.Grid(d)
    g.upload(h)
    gdt, _ = g.compute_dt(0, 0.0, math.inf, 0.5)
    assert gdt == dt
    ora.sweep(axis, dt)
    g.sweep(axis, dt)
    check(g.download(), ora.grid, f"allaire axis={axis} ws={ws}")


@pytest.mark.parametrize("splitting", [cfdart.NAIVE, cfdart.STRANG, cfdart.ALTERNATING])
def test_euler_run_splittings(gpu, splitting):
    d, h = euler_small(96, 80, (1, 0, 1, 0), cfdart.VANLEER, seed=11)
    print(f"case euler run splitting={splitting}", flush=True)
    ora = oracle.Sim(d, h)
    k, hist = ora.run(splitting, 30, cfl=0.8)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n = g.run(splitting, 30, cfl=0.8)
    assert n == 30 and t == ora.t
    check(g.download(), ora.grid, f"euler run split={splitting}")
    assert g.stats()[:3] == ora.counters()[:3]


def test_euler_200_to_T(gpu):
    """Config euler_200 in full: run to t = 0.2 with the reference's loop (dt clipped to T)."""
    d, h = cfdart.configs.euler_200()
    ora = oracle.Sim(d, h)
    k, _ = ora.run(cfdart.ALTERNATING, 10000, T=0.2, cfl=0.8)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n = g.run(cfdart.ALTERNATING, 10000, T=0.2, cfl=0.8)
    print(f"euler_200: steps gpu={n} oracle={k} t={t}", flush=True)
    assert n == k and t == ora.t == 0.2
    check(g.download(), ora.grid, "euler_200 t=0.2")


def test_allaire_run(gpu):
    d, h = cfdart.configs.allaire_bubbles(96, 6, seed=2, ny=64)
    ora = oracle.Sim(d, h)
    ora.run(cfdart.ALTERNATING, 40, cfl=0.5)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n = g.run(cfdart.ALTERNATING, 40, cfl=0.5)
    assert n == 40 and t == ora.t
    check(g.download(), ora.grid, "allaire 96x64 40 steps")
