"""GPU parity at the benchmark's full sizes.

euler_2048 (2048x2048, 64 discs) after 100 alternating steps at CFL 0.8, allaire_1024
(1024x1024, 16 bubbles behind a planar shock) after a 100-step prefix and the 2048x256 slab of
euler_2048 after 30 steps, all against the oracle on the same seeded initial state: bitwise,
and per-field relative L1/Linf within 1e-10. Also size-independent checks on the device: the
whole 2000-step allaire_1024 run agrees across tile shapes and keeps every state physical,
mass and energy are conserved in a closed (all-reflective) 2048x2048 box over 100 steps, and a
free stream is preserved.
"""
import math
import time

import numpy as np
import pytest

import cfdart
import oracle
from util import bitwise_equal, field_errors, interior, n_bit_diffs

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _compare(a, b, label):
    a, b = interior(a), interior(b)
    errs = field_errors(a, b)
    print(f"{label}: bit-diffs={n_bit_diffs(a, b)} relL1/Linf per field={errs}", flush=True)
    for l1, li in errs:
        assert l1 <= TOL and li <= TOL, (label, errs)
    assert bitwise_equal(a, b), label


def test_euler_2048_100_steps(gpu):
    d, h = cfdart.configs.euler_2048()
    print("case euler_2048, 100 alternating steps, CFL 0.8", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    tg, ng = g.run(cfdart.ALTERNATING, 100, cfl=0.8)
    t0 = time.perf_counter()
    ora = oracle.Sim(d, h)
    no, _ = ora.run(cfdart.ALTERNATING, 100, cfl=0.8)
    print(f"oracle: {time.perf_counter() - t0:.1f} s", flush=True)
    assert ng == no == 100 and tg == ora.t
    _compare(g.download(), ora.grid, "euler_2048 n=100")
    assert g.stats() == ora.counters()


def test_allaire_1024_prefix(gpu):
    d, h = cfdart.configs.allaire_1024()
    print("case allaire_1024, 100 alternating steps, CFL 0.5", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    tg, ng = g.run(cfdart.ALTERNATING, 100, cfl=0.5)
    ora = oracle.Sim(d, h)
    no, _ = ora.run(cfdart.ALTERNATING, 100, cfl=0.5)
    assert ng == no == 100 and tg == ora.t
    _compare(g.download(), ora.grid, "allaire_1024 n=100")
    assert g.stats() == ora.counters()


def test_allaire_1024_full_run_tiles_agree(gpu):
    """The whole 2000-step allaire_1024 run (too long for the oracle in a test): the transposed
    8-column tile and 8-row strips end on the same bits with the same counters, no state fails
    the physical-state test, and z stays inside its clip bounds."""
    d, h = cfdart.configs.allaire_1024()
    out = []
    for tile in [(cfdart.Y_TRANSPOSED_NARROW, 2), (8, 4)]:
        print(f"case allaire_1024, 2000 steps, tile {tile}", flush=True)
        g = cfdart.Grid(d)
        g.set_tile(*tile)
        g.upload(h)
        t, n = g.run(cfdart.ALTERNATING, 2000, cfl=0.5)
        z = interior(g.download())[..., 5]
        out.append((t, n, g.checksum(), g.stats()))
        g.close()
        assert n == 2000 and math.isfinite(t)
        assert z.min() >= d.z_lo and z.max() <= d.z_hi
    assert out[0] == out[1]
    assert out[0][3][3] == 0  # no unphysical state


def test_euler_slab_parity(gpu):
    """One GPU's 2048x256 share of euler_2048 (the sharding measurement's grid), 30 steps."""
    d, h = cfdart.configs.euler_slab()
    print("case euler_slab 2048x256, 30 alternating steps", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    tg, ng = g.run(cfdart.ALTERNATING, 30, cfl=0.8)
    ora = oracle.Sim(d, h)
    no, _ = ora.run(cfdart.ALTERNATING, 30, cfl=0.8)
    assert ng == no == 30 and tg == ora.t
    _compare(g.download(), ora.grid, "euler_slab n=30")
    assert g.stats() == ora.counters()


def test_conservation_closed_box_2048(gpu):
    d, h = cfdart.configs.euler_discs(2048, 64, (0.005, 0.03), (1, 1, 1, 1), seed=1)
    print("case closed box 2048^2, 100 steps", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 100, cfl=0.8)
    s = g.download()
    assert g.stats()[2] == 0  # no positivity repair, which would not conserve
    for k in (0, 3):  # mass and energy; momentum is not conserved against walls
        # exactly rounded sums: a plain float sum of 4M terms carries ~1e-13 error of its own
        tot0 = math.fsum(interior(h)[..., k].ravel())
        tot = math.fsum(interior(s)[..., k].ravel())
        rel = abs(tot - tot0) / abs(tot0)
        print(f"component {k}: relative change {rel:.3e}", flush=True)
        assert rel <= 1e-13


def test_free_stream_2048(gpu):
    d, _ = cfdart.configs.euler_2048()
    d.bc[0] = d.bc[1] = d.bc[2] = d.bc[3] = 0
    h = np.empty((d.ny + 4, d.nx + 4, 4))
    h[...] = (1.3, 1.3 * 0.7, 1.3 * -0.4, 2.5 / 0.4 + 0.5 * 1.3 * (0.49 + 0.16))
    print("case free stream 2048^2, 20 steps", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 20, cfl=0.8)
    assert bitwise_equal(interior(g.download()), interior(h))
