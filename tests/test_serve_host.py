"""Host side of serving (no GPU): the request trace and the request -> grid mapping."""
import math

import numpy as np
import pytest

import cfdart
from cfdart import serve


def test_trace_is_seeded_and_has_the_stated_distribution():
    a = serve.trace(seed=11, rate=50.0, duration=120.0)
    b = serve.trace(seed=11, rate=50.0, duration=120.0)
    assert len(a) == len(b) and all(x.arrival == y.arrival and x.n == y.n and x.seed == y.seed for x, y in zip(a, b))
    arr = np.array([r.arrival for r in a])
    n = np.array([r.n for r in a], dtype=float)
    steps = np.array([r.steps for r in a])
    model = np.array([r.model for r in a])
    # Poisson(50/s) over 120 s: 6000 +- 3 sigma
    assert abs(len(a) - 6000) < 3 * math.sqrt(6000)
    assert np.all(np.diff(arr) >= 0) and arr[-1] <= 120.0
    gaps = np.diff(np.concatenate(([0.0], arr)))
    assert abs(gaps.mean() - 0.02) < 0.002
    assert abs((model == 0).mean() - 0.7) < 0.03
    assert n.min() >= 64 and n.max() <= 2048 and steps.min() >= 50 and steps.max() <= 2000
    # log-uniform side: ln n uniform on [ln 64, ln 2048]
    u = (np.log(n) - math.log(64)) / (math.log(2048) - math.log(64))
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
    assert abs(steps.mean() - 1025) < 25
    assert len({r.seed for r in a}) == len(a)


def test_request_setup_matches_the_named_workloads():
    r = serve.requests([(0.0, 0, 96, 100, 5), (0.0, 1, 80, 100, 6)])
    d, p, cfl = serve.setup(r[0])
    d2, p2 = cfdart.configs.euler_discs_params(96, 8, (0.03, 0.10), (0, 0, 0, 0), seed=5)
    assert cfl == 0.8 and isinstance(p, cfdart.DiscParams)
    assert bytes(d) == bytes(d2) and bytes(p) == bytes(p2)
    d, p, cfl = serve.setup(r[1])
    d2, p2 = cfdart.configs.allaire_bubbles_params(80, 16, seed=6)
    assert cfl == 0.5 and isinstance(p, cfdart.BubbleParams)
    assert bytes(d) == bytes(d2) and bytes(p) == bytes(p2)


def test_state_checksum_host_restatement():
    # position-sensitive and value-sensitive; independent of the ghost frame
    s = np.random.default_rng(0).standard_normal((10, 12, 4))
    c = cfdart.state_checksum(s)
    t = s.copy()
    t[0, 0, 0] = 99.0
    assert cfdart.state_checksum(t) == c
    t = s.copy()
    t[3, 4, 1] = np.nextafter(t[3, 4, 1], np.inf)
    assert cfdart.state_checksum(t) != c
    t = s.copy()
    t[[3, 4], 5, 2] = t[[4, 3], 5, 2]
    assert cfdart.state_checksum(t) != c


def test_serve_argument_errors_need_no_gpu():
    L = serve._bind()
    import ctypes as C
    out = (serve.Request * 4)()
    n = C.c_int()
    assert L.cfdart_serve_trace(1, -1.0, 1.0, 0.5, 64, 128, 1, 2, out, 4, C.byref(n)) == -1   # rate
    assert L.cfdart_serve_trace(1, 1.0, 1.0, 0.5, 128, 64, 1, 2, out, 4, C.byref(n)) == -1    # n_hi < n_lo
    assert L.cfdart_serve_trace(1, 1.0, 1.0, 0.5, 64, 128, 5, 2, out, 4, C.byref(n)) == -1    # steps_hi < steps_lo
    bad = serve.requests([(0.0, 7, 64, 10, 1)])
    with pytest.raises(cfdart.CfdartError):
        serve.setup(bad[0])  # unknown model
    unsorted = serve.requests([(1.0, 0, 64, 10, 1), (0.5, 0, 64, 10, 2)])
    res = (serve.RequestResult * 2)()
    s = serve.ServeSummary()
    o = serve.opts()
    assert L.cfdart_serve_batched(unsorted, 2, C.byref(o), res, C.byref(s)) == -1           # arrivals not sorted
    o2 = serve.opts(chunk_steps=7, frame_every=50)
    ok = serve.requests([(0.0, 0, 64, 10, 1)])
    assert L.cfdart_serve_batched(ok, 1, C.byref(o2), res, C.byref(s)) == -1                # chunk must divide frames
