"""GPU: device-side initial conditions, state checksums, and the serving loop.

* The device generators write bitwise the state of the host generators.
* The device checksum equals its host restatement on the downloaded state.
* Served requests (continuous batching, requests admitted while others run, under a memory
  budget) end bitwise where the same requests end alone with the plain solver, and where the
  oracle ends for a request recomputed on the CPU.
"""
import numpy as np
import pytest

import cfdart
import oracle
from cfdart import serve
from util import bitwise_equal, interior

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model,n,ny,k", [(0, 96, 96, 8), (0, 130, 77, 64), (0, 2048, 2048, 64),
                                          (1, 100, 100, 16), (1, 150, 90, 16), (1, 1024, 1024, 16)])
def test_device_ic_equals_host_ic(gpu, model, n, ny, k):
    print(f"case device IC model={model} {n}x{ny} k={k}", flush=True)
    if model == 0:
        d, p = cfdart.configs.euler_discs_params(n, k, (0.005, 0.1), (0, 0, 0, 0), seed=n + k, ny=ny)
        h = cfdart.random_discs(d, p)
    else:
        d, p = cfdart.configs.allaire_bubbles_params(n, k, seed=n + k, ny=ny)
        h = cfdart.shocked_bubbles(d, p)
    g = cfdart.Grid(d)
    g.init_generated(p)
    s = g.download()
    assert bitwise_equal(interior(s), interior(h))
    assert g.checksum() == cfdart.state_checksum(h)


def test_batch_generated_equals_uploaded(gpu):
    d, p = cfdart.configs.allaire_bubbles_params(120, 16, seed=4, ny=70)
    h = cfdart.shocked_bubbles(d, p)
    b = cfdart.Batch(4)
    s1 = b.add_generated(d, p, max_steps=15, cfl=0.5)
    s2 = b.add(d, h, max_steps=15, cfl=0.5)
    b.step(20)
    b.sync()
    steps, flags = b.poll(4)
    assert steps[s1] == steps[s2] == 15 and flags[s1] == flags[s2] == 0
    assert b.checksum(s1) == b.checksum(s2)
    assert bitwise_equal(interior(b.download(s1)), interior(b.download(s2)))


def _mixed_requests():
    rng = np.random.default_rng(3)
    items = []
    t = 0.0
    for i in range(24):
        t += rng.exponential(0.01)
        model = int(rng.uniform() < 0.6)
        n = int(round(np.exp(rng.uniform(np.log(48), np.log(256)))))
        items.append((t, 1 - model, n, int(rng.integers(20, 121)), 500 + i))
    return serve.requests(items)


def test_served_equals_alone(gpu):
    reqs = _mixed_requests()
    print(f"case serve {len(reqs)} requests, realtime", flush=True)
    o = serve.opts(capacity=64, chunk_steps=10, frame_every=20, frame_host_bytes=64 << 20, realtime=True)
    rb, sb = serve.serve_batched(reqs, o)
    ro, so = serve.serve_one_at_a_time(reqs, o)
    assert sb.max_live > 1  # requests overlapped on the device
    for q, a, b in zip(reqs, rb, ro):
        assert a.status == 0 and b.status == 0
        assert a.steps_done == b.steps_done == q.steps
        assert a.frames == b.frames == q.steps // 20
        assert a.checksum == b.checksum, (q.model, q.n, q.steps)
        assert a.finished >= q.arrival
    assert sb.cell_updates == so.cell_updates == sum(q.n * q.n * q.steps for q in reqs)
    # one request end to end against the oracle
    q = reqs[0]
    d, p, cfl = serve.setup(q)
    h = cfdart.random_discs(d, p) if q.model == 0 else cfdart.shocked_bubbles(d, p)
    s = oracle.Sim(d, h)
    k, _ = s.run(cfdart.ALTERNATING, q.steps, cfl=cfl)
    assert k == q.steps
    assert cfdart.state_checksum(s.grid) == rb[0].checksum


def test_served_odd_chunks(gpu):
    """Chunks of 5 steps admit requests at odd batch steps, so the live grids no longer share a
    step parity and the batch launches both axes in both phases: every result is still bitwise
    its one-at-a-time run."""
    reqs = _mixed_requests()
    print(f"case serve {len(reqs)} requests, chunks of 5 steps", flush=True)
    # a live-cell cap staggers the admissions over the run (at multiples of 5 steps)
    o = serve.opts(capacity=64, chunk_steps=5, frame_every=0, realtime=False, live_cells=3 * 160 * 160)
    rb, sb = serve.serve_batched(reqs, o)
    ro, _ = serve.serve_one_at_a_time(reqs, o)
    assert sb.max_live > 1
    for q, a, b in zip(reqs, rb, ro):
        assert a.status == 0 and b.status == 0 and a.steps_done == b.steps_done == q.steps
        assert a.checksum == b.checksum, (q.model, q.n, q.steps)


def test_admission_waits_for_memory(gpu):
    reqs = _mixed_requests()
    # room for the grid state of about three of these requests at a time
    o = serve.opts(mem_budget=3 * 2 * 6 * 8 * 256 * 256 + (16 << 20), capacity=64, chunk_steps=10, frame_every=0,
                   realtime=False)
    rb, sb = serve.serve_batched(reqs, o)
    ro, _ = serve.serve_one_at_a_time(reqs, o)
    print(f"max live {sb.max_live}, peak bytes {sb.peak_state_bytes:.3g}", flush=True)
    assert 1 < sb.max_live < len(reqs)
    assert sb.peak_state_bytes <= o.mem_budget
    for a, b in zip(rb, ro):
        assert a.status == 0 and a.checksum == b.checksum


def test_request_larger_than_budget_is_refused(gpu):
    reqs = serve.requests([(0.0, 0, 64, 20, 1), (0.0, 0, 512, 20, 2), (0.0, 1, 64, 20, 3)])
    o = serve.opts(mem_budget=8 << 20, capacity=8, chunk_steps=10, frame_every=0, realtime=False)
    rb, _ = serve.serve_batched(reqs, o)
    assert rb[0].status == 0 and rb[2].status == 0
    assert rb[1].status == -2 and rb[1].steps_done == 0


def test_live_cell_cap(gpu):
    """Admission capped by live cells: the cap holds whenever more than one grid is live, every
    request still completes, and the results are bitwise those of the uncapped server."""
    reqs = _mixed_requests()
    cap = 3 * 160 * 160
    o = serve.opts(capacity=64, chunk_steps=10, frame_every=0, realtime=False)
    oc = serve.opts(capacity=64, chunk_steps=10, frame_every=0, realtime=False, live_cells=cap)
    rb, sb = serve.serve_batched(reqs, o)
    rc, sc = serve.serve_batched(reqs, oc)
    print(f"max live uncapped {sb.max_live}, capped {sc.max_live}", flush=True)
    assert sc.max_live < sb.max_live
    for a, b in zip(rb, rc):
        assert a.status == 0 and b.status == 0 and a.checksum == b.checksum
    # every admission happened at a moment when the grids live then, plus the new one, fit the cap
    # (or the new one was alone)
    ivals = sorted((r.admitted, r.finished, q.n * q.n) for q, r in zip(reqs, rc))
    for t, _, cells in ivals:
        live = [c for a, f, c in ivals if a < t < f]
        assert not live or sum(live) + cells <= cap
