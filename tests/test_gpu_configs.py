"""The batched and served configurations at their configured workloads, each result against the
same grid solved alone with the plain single-grid solver:

* batched: 256 grids (192 Euler, 64 Allaire), sides drawn independently log-uniform in
  [128, 1024], 500 steps each, in both arithmetic modes; every grid's final state has the
  checksum of its lone run (cfdart_checksum hashes every interior value's bits);
* served: the configured request stream (Poisson arrivals at 50/s, 70 % Euler, sides log-uniform
  in [64, 2048], steps uniform in [50, 2000], a frame every 50 steps, 200 GB budget, at most 32 M
  live cells) replayed in real time over its first 2 s of arrivals, every result's checksum equal
  to its one-at-a-time run, with requests overlapping on the device."""
import numpy as np
import pytest

import cfdart
from cfdart import serve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("arith", [cfdart.ARITH_EXACT, cfdart.ARITH_FAST], ids=["exact", "fast"])
def test_batched_config_equals_alone(gpu, arith):
    cases = cfdart.configs.batched_mix(256)
    steps = 500
    assert sum(1 for c in cases if c[0].model == cfdart.EULER) == 192
    print(f"case batched_mix(256) x {steps} steps, arith {arith}, "
          f"{sum(c[0].nx * c[0].ny for c in cases) / 1e6:.1f} M cells", flush=True)
    b = cfdart.Batch(len(cases))
    b.set_arith(arith)
    slots = [b.add(d, h, splitting=cfdart.ALTERNATING, max_steps=steps, cfl=cfl) for d, h, cfl in cases]
    b.step(steps)
    b.sync()
    done, flags = b.poll(len(cases))
    assert all(done[s] == steps and flags[s] == 0 for s in slots)
    batched = [b.checksum(s) for s in slots]
    b.close()
    same = 0
    for (d, h, cfl), want in zip(cases, batched):
        g = cfdart.Grid(d)
        g.set_arith(arith)
        g.upload(h)
        t, n = g.run(cfdart.ALTERNATING, steps, cfl=cfl)
        assert n == steps
        same += int(g.checksum() == want)
        g.close()
    assert same == len(cases)


def test_served_config_equals_alone(gpu):
    reqs = serve.trace(seed=7, rate=50.0, duration=2.0)
    work = sum(q.n * q.n * q.steps for q in reqs)
    print(f"case served: {len(reqs)} requests over 2 s, {work / 1e9:.1f} G cell-updates", flush=True)
    assert len(reqs) > 50
    sides = np.array([q.n for q in reqs])
    assert sides.min() >= 64 and sides.max() <= 2048
    o = serve.opts(mem_budget=200 * 10**9, capacity=8192, chunk_steps=10, frame_every=50, frame_host_bytes=1 << 30,
                   realtime=True, live_cells=32 * 10**6)
    rb, sb = serve.serve_batched(reqs, o)
    ro, so = serve.serve_one_at_a_time(reqs, o)
    print(f"served: makespan {sb.makespan:.2f} s, max live {sb.max_live}, frames {sb.frames}; "
          f"one at a time {so.makespan:.2f} s", flush=True)
    assert sb.max_live > 1
    for q, a, b in zip(reqs, rb, ro):
        assert a.status == 0 and b.status == 0
        assert a.steps_done == b.steps_done == q.steps
        assert a.frames == b.frames == q.steps // 50
        assert a.checksum == b.checksum, (q.model, q.n, q.steps)
    assert sb.cell_updates == so.cell_updates == work
