"""GPU accuracy of the fast arithmetic (cfdart_set_arith with CFDART_ARITH_FAST): divisions are
the numerator times the divisor's reciprocal refined by one Newton step, square roots take one
Goldschmidt step, products may fuse into sums, and work that only reproduces the exact
arithmetic's roundings is skipped. The results are not bitwise the reference's arithmetic; the
bar is per-field relative L1 and Linf within TOL = 1e-10 of the oracle on the same seeded
inputs: euler_2048 and allaire_1024 after 100 alternating steps, the first-order Godunov scheme,
every tile shape on a ragged grid and every splitting. Mass and energy stay conserved to round-off in a closed box, a
free stream is preserved bitwise, and the exact mode on the same handle is bitwise again. A batch
and the serving loop in the fast arithmetic give bitwise the lone fast runs."""
import math

import numpy as np
import pytest

import cfdart
import oracle
from test_gpu_shapes import state
from util import bitwise_equal, field_errors, interior

pytestmark = pytest.mark.gpu
TOL = 1e-10
FAST = cfdart.ARITH_FAST


def _within(a, b, label, ng=2, momentum_norm=False):
    """Per-field relative L1 and Linf of a against b. With momentum_norm, the two momentum
    components are measured against the momentum vector's L1 and Linf norms instead of their
    own: a component that is zero in exact arithmetic (v = 0 of a planar shock that has not yet
    met a bubble) holds only round-off, against which any other rounding is relatively large."""
    a, b = interior(a, ng), interior(b, ng)
    errs = field_errors(a, b)
    if momentum_norm:
        mom = (1, 2) if a.shape[-1] == 4 else (2, 3)
        l1_den = sum(float(np.abs(b[..., k]).sum()) for k in mom)
        li_den = max(float(np.abs(b[..., k]).max()) for k in mom)
        for k in mom:
            d = np.abs(a[..., k] - b[..., k])
            errs[k] = (float(d.sum()) / l1_den, float(d.max()) / li_den)
    print(f"{label}: relL1/Linf per field={errs}", flush=True)
    for l1, li in errs:
        assert l1 <= TOL and li <= TOL, (label, errs)
    return errs


def _run_pair(d, h, splitting, steps, cfl, tile=None):
    g = cfdart.Grid(d)
    if tile:
        g.set_tile(*tile)
    g.set_arith(FAST)
    g.upload(h)
    tg, ng = g.run(splitting, steps, cfl=cfl)
    ora = oracle.Sim(d, h)
    no, _ = ora.run(splitting, steps, cfl=cfl)
    assert ng == no == steps
    assert abs(tg - ora.t) <= 1e-12 * ora.t, (tg, ora.t)
    return g, ora


def test_euler_2048_100_steps(gpu):
    d, h = cfdart.configs.euler_2048()
    print("case fast arithmetic, euler_2048, 100 alternating steps", flush=True)
    g, ora = _run_pair(d, h, cfdart.ALTERNATING, 100, 0.8)
    errs = _within(g.download(), ora.grid, "euler_2048 fast n=100")
    assert max(l1 for l1, _ in errs) > 0.0  # the fast arithmetic is in use (not bitwise)
    print(f"counters gpu {g.stats()} oracle {ora.counters()}", flush=True)


def test_allaire_1024_100_steps(gpu):
    d, h = cfdart.configs.allaire_1024()
    print("case fast arithmetic, allaire_1024, 100 alternating steps", flush=True)
    g, ora = _run_pair(d, h, cfdart.ALTERNATING, 100, 0.5)
    _within(g.download(), ora.grid, "allaire_1024 fast n=100")
    assert g.stats()[3] == 0  # no unphysical state


@pytest.mark.parametrize("model", [0, 1])
def test_godunov(gpu, model):
    d, h = state(model, 200, 130, cfdart.GODUNOV, seed=11 + model)
    print(f"case fast arithmetic, Godunov model={model}", flush=True)
    g, ora = _run_pair(d, h, cfdart.ALTERNATING, 40, 0.8 if model == 0 else 0.5)
    _within(g.download(), ora.grid, f"godunov model={model}", ng=d.ng)


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("tile", [(cfdart.Y_TRANSPOSED, 4), (cfdart.Y_TRANSPOSED_NARROW, 2), (12, 3)])
@pytest.mark.parametrize("splitting", [cfdart.NAIVE, cfdart.STRANG, cfdart.ALTERNATING])
def test_tiles_and_splittings(gpu, model, tile, splitting):
    d, h = state(model, 150, 97, cfdart.MUSCL_HANCOCK, seed=41 + model)
    print(f"case fast arithmetic model={model} tile={tile} splitting={splitting}", flush=True)
    g, ora = _run_pair(d, h, splitting, 12, 0.8 if model == 0 else 0.5, tile=tile)
    _within(g.download(), ora.grid, f"model={model} tile={tile} splitting={splitting}")


@pytest.mark.parametrize("model", [0, 1])
def test_chunked_run_equals_one_run(gpu, model):
    """A run split into several cfdart_run calls (each starts from a freshly reduced dt) ends on
    the same bits as one run: the first step's wave speeds take the same arithmetic as the last
    sweep's."""
    d, h = state(model, 180, 120, cfdart.MUSCL_HANCOCK, seed=23 + model)
    cfl = 0.8 if model == 0 else 0.5
    print(f"case fast arithmetic chunked run model={model}", flush=True)
    out = []
    for chunks in ([40], [10, 10, 15, 5]):
        g = cfdart.Grid(d)
        g.set_arith(FAST)
        g.upload(h)
        for k in chunks:
            t, n = g.run(cfdart.ALTERNATING, k, cfl=cfl)
        out.append((t, n, g.checksum()))
        g.close()
    assert out[0] == out[1] and out[0][1] == 40


def test_conservation_closed_box(gpu):
    d, h = cfdart.configs.euler_discs(1024, 32, (0.005, 0.03), (1, 1, 1, 1), seed=3)
    print("case fast arithmetic, closed box 1024^2, 100 steps", flush=True)
    g = cfdart.Grid(d)
    g.set_arith(FAST)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 100, cfl=0.8)
    s = g.download()
    assert g.stats()[2] == 0  # no positivity repair, which would not conserve
    for k in (0, 3):
        tot0 = math.fsum(interior(h)[..., k].ravel())
        tot = math.fsum(interior(s)[..., k].ravel())
        rel = abs(tot - tot0) / abs(tot0)
        print(f"component {k}: relative change {rel:.3e}", flush=True)
        assert rel <= 1e-13


def test_free_stream(gpu):
    d, _ = cfdart.configs.euler_discs(512, 4, (0.05, 0.1), (0, 0, 0, 0), seed=1)
    h = np.empty((d.ny + 4, d.nx + 4, 4))
    h[...] = (1.3, 1.3 * 0.7, 1.3 * -0.4, 2.5 / 0.4 + 0.5 * 1.3 * (0.49 + 0.16))
    print("case fast arithmetic, free stream 512^2, 20 steps", flush=True)
    g = cfdart.Grid(d)
    g.set_arith(FAST)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 20, cfl=0.8)
    assert bitwise_equal(interior(g.download()), interior(h))


def _riemann_x(nx, ny, left, right):
    """Euler states (rho, u, p) left and right of x = 1/2, transmissive boundaries."""
    d = cfdart.make_desc(cfdart.EULER, nx, ny, lx=1.0, ly=ny / nx, gamma1=1.4, bc=(0, 0, 0, 0),
                         limiter=cfdart.VANLEER)
    h = np.zeros((ny + 4, nx + 4, 4))
    for j in range(nx):
        rho, u, p = left if (j + 0.5) / nx < 0.5 else right
        h[2:-2, 2 + j] = (rho, rho * u, 0.0, p / 0.4 + 0.5 * rho * u * u)
    return d, h


def test_evolve_fallbacks(gpu):
    """A near-vacuum double rarefaction (rho 1e-3, u -20, p 1e-6 against rho 1, u 20, p 1): the
    evolved half-step states leave the physical set and the faces fall back to the cells' own
    states (the oracle counts 368 such faces in 60 steps). The exact arithmetic gives the
    oracle's bits and counters; the fast one takes its positivity decisions on the half-step
    states' primitives and stays within TOL."""
    d, h = _riemann_x(128, 8, (1e-3, -20.0, 1e-6), (1.0, 20.0, 1.0))
    ora = oracle.Sim(d, h)
    ora.run(cfdart.ALTERNATING, 60, cfl=0.8)
    assert ora.counters()[1] > 0
    for mode in (cfdart.ARITH_EXACT, FAST):
        print(f"case evolve fallbacks, arith={mode}", flush=True)
        g = cfdart.Grid(d)
        g.set_arith(mode)
        g.upload(h)
        t, n = g.run(cfdart.ALTERNATING, 60, cfl=0.8)
        print(f"counters gpu {g.stats()} oracle {ora.counters()}", flush=True)
        if mode == FAST:
            assert n == 60 and abs(t - ora.t) <= 1e-12 * ora.t
            _within(g.download(), ora.grid, "evolve fallbacks fast")
            assert g.stats()[1] > 0
        else:
            assert n == 60 and t == ora.t
            assert bitwise_equal(interior(g.download()), interior(ora.grid))
            assert g.stats() == ora.counters()
        g.close()


def test_exact_again_on_the_same_handle(gpu):
    d, h = state(0, 130, 70, cfdart.MUSCL_HANCOCK, seed=5)
    print("case fast then exact arithmetic on one handle", flush=True)
    g = cfdart.Grid(d)
    g.set_arith(FAST)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 3, cfl=0.8)
    g.set_arith(cfdart.ARITH_EXACT)
    g.upload(h)
    t, n = g.run(cfdart.ALTERNATING, 6, cfl=0.8)
    ora = oracle.Sim(d, h)
    ora.run(cfdart.ALTERNATING, 6, cfl=0.8)
    assert n == 6 and t == ora.t
    assert bitwise_equal(interior(g.download()), interior(ora.grid))
    assert g.stats() == ora.counters()
    with pytest.raises(cfdart.CfdartError):
        g.set_arith(2)


def test_batch_fast_equals_alone_fast_and_oracle(gpu):
    """A mixed batch in the fast arithmetic: every grid bitwise equal to the same grid run alone in
    the fast arithmetic (the same arithmetic on other tiles), and within TOL of the oracle; the
    batch switched back to the exact arithmetic is bitwise the exact lone runs again."""
    from test_gpu_batch import mixed_cases
    cases = mixed_cases(7)
    for mode in (FAST, cfdart.ARITH_EXACT):
        b = cfdart.Batch(16)
        b.set_arith(FAST)
        b.set_arith(mode)
        slots = [b.add(d, h, max_steps=steps, cfl=cfl, y_strip_rows=cfdart.Y_TRANSPOSED, x_rows_per_wave=2)
                 for d, h, cfl, steps in cases]
        b.step(30)
        b.sync()
        for (d, h, cfl, steps), s in zip(cases, slots):
            print(f"case batch arith={mode} slot={s} model={d.model} {d.nx}x{d.ny} steps={steps}", flush=True)
            g = cfdart.Grid(d)
            g.set_arith(mode)
            g.upload(h)
            t, n = g.run(cfdart.ALTERNATING, steps, cfl=cfl)
            tb, nb, act = b.status(s)
            assert nb == n == steps and tb == t and not act
            assert bitwise_equal(interior(b.download(s)), interior(g.download()))
            ora = oracle.Sim(d, h)
            ora.run(cfdart.ALTERNATING, steps, cfl=cfl)
            if mode == FAST:
                _within(g.download(), ora.grid, f"batch fast slot={s}", momentum_norm=True)
            else:
                assert bitwise_equal(interior(g.download()), interior(ora.grid))
            g.close()
        b.close()


def test_served_fast_equals_alone(gpu):
    """The serving loop in the fast arithmetic: every request's result (checksum) equals its
    one-at-a-time run in the fast arithmetic, with frames (so the lone runs go in chunks)."""
    from cfdart import serve
    from test_gpu_serve import _mixed_requests
    reqs = _mixed_requests()
    # frames every 20 steps: the one-at-a-time runs go in chunks of 20 steps, the batch does not
    o = serve.opts(capacity=64, chunk_steps=10, frame_every=20, frame_host_bytes=64 << 20, realtime=True, arith=FAST)
    print(f"case served fast, {len(reqs)} requests", flush=True)
    rb, sb = serve.serve_batched(reqs, o)
    ro, so = serve.serve_one_at_a_time(reqs, o)
    assert all(r.status == 0 for r in rb) and all(r.status == 0 for r in ro)
    assert [r.checksum for r in rb] == [r.checksum for r in ro]
    assert sb.max_live > 1
