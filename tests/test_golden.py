"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py from the
oracle): the oracle must reproduce them bitwise; the GPU path too."""
import glob
import os

import numpy as np
import pytest

import cfdart
import oracle
from util import bitwise_equal, interior

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def desc_from_array(a):
    d = cfdart.Desc()
    (d.model, d.nx, d.ny, d.ng) = (int(a[0]), int(a[1]), int(a[2]), int(a[3]))
    d.x0, d.y0, d.dx, d.dy, d.gamma1, d.pinf1, d.gamma2, d.pinf2 = a[4:12]
    for i in range(4):
        d.bc[i] = int(a[12 + i])
    d.limiter, d.wave_speeds, d.colsweep_uses_dx = int(a[16]), int(a[17]), int(a[18])
    d.z_lo, d.z_hi = a[19], a[20]
    d.flux_solver = int(a[21]) if len(a) > 21 else 0  # fixtures written before the field existed
    return d


def load(path):
    f = np.load(path, allow_pickle=False)
    return (desc_from_array(f["desc"]), f["initial"], f["final"], int(f["splitting"]), int(f["steps"]),
            float(f["cfl"]), f["dt"], f["counters"])


def test_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    d, h0, h1, split, steps, cfl, dts, cnt = load(path)
    s = oracle.Sim(d, h0)
    k, hist = s.run(split, steps, cfl=cfl)
    assert k == steps
    assert np.array_equal(hist, dts)
    g = d.ng
    assert bitwise_equal(s.grid[g:-g, g:-g], h1[g:-g, g:-g])
    assert list(cnt) == s.counters()


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_reproduces_fixture(gpu, path):
    d, h0, h1, split, steps, cfl, dts, cnt = load(path)
    print(f"case golden {os.path.basename(path)}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h0)
    t, n = g.run(split, steps, cfl=cfl)
    assert n == steps
    ng = d.ng
    assert bitwise_equal(g.download()[ng:-ng, ng:-ng], h1[ng:-ng, ng:-ng])
