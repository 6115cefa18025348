"""Batched execution: every grid in a mixed Euler/Allaire batch of uneven sizes gives bitwise
the result of the same grid run alone through the plain solver (and hence the oracle), including
grids that join late, finish early or are removed; admission respects the memory budget."""
import math

import numpy as np
import pytest

import cfdart
import oracle
from util import bitwise_equal, interior

pytestmark = pytest.mark.gpu


def mixed_cases(seed=0, n=8):
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(n):
        nx, ny = int(rng.integers(24, 130)), int(rng.integers(24, 130))
        if k % 4 == 3:
            d, h = cfdart.configs.allaire_bubbles(nx, 2, seed=100 + k, ny=ny)
            cfl = 0.5
        else:
            bc = tuple(int(b) for b in rng.integers(0, 2, 4))
            d, h = cfdart.configs.euler_discs(nx, 3, (0.05, 0.2), bc, seed=100 + k, ny=ny)
            cfl = 0.8
        cases.append((d, h, cfl, int(rng.integers(5, 25))))
    return cases


def alone(d, h, cfl, steps, splitting=cfdart.ALTERNATING):
    g = cfdart.Grid(d)
    g.upload(h)
    t, n = g.run(splitting, steps, cfl=cfl)
    return g.download(), t, n


@pytest.mark.parametrize("splitting", [cfdart.ALTERNATING, cfdart.NAIVE, cfdart.STRANG])
def test_batch_equals_alone(gpu, splitting):
    cases = mixed_cases(1)
    b = cfdart.Batch(16)
    slots = []
    for d, h, cfl, steps in cases:
        slots.append(b.add(d, h, splitting=splitting, max_steps=steps, cfl=cfl, y_strip_rows=cfdart.Y_TRANSPOSED,
                           x_rows_per_wave=2))
    b.step(30)
    b.sync()
    for (d, h, cfl, steps), s in zip(cases, slots):
        ref, t, n = alone(d, h, cfl, steps, splitting)
        tb, nb, act = b.status(s)
        print(f"case batch slot={s} model={d.model} {d.nx}x{d.ny} steps={steps}", flush=True)
        assert nb == n == steps and tb == t and not act
        assert bitwise_equal(interior(b.download(s)), interior(ref))
    b.close()


def test_batch_join_late_and_remove(gpu):
    cases = mixed_cases(2, 6)
    b = cfdart.Batch(8)
    s0 = b.add(*cases[0][:2], max_steps=cases[0][3], cfl=cases[0][2])
    s1 = b.add(*cases[1][:2], max_steps=40, cfl=cases[1][2])
    b.step(3)
    s2 = b.add(*cases[2][:2], max_steps=cases[2][3], cfl=cases[2][2])  # joins at global step 3
    b.step(2)
    b.remove(s1)  # leaves mid-run
    s3 = b.add(*cases[3][:2], max_steps=cases[3][3], cfl=cases[3][2])  # may reuse s1's memory
    b.step(40)
    b.sync()
    for s, c in ((s0, cases[0]), (s2, cases[2]), (s3, cases[3])):
        d, h, cfl, steps = c
        ref, t, n = alone(d, h, cfl, steps)
        assert b.status(s)[1] == steps
        assert bitwise_equal(interior(b.download(s)), interior(ref))


def test_batch_matches_oracle_run_to_T(gpu):
    d, h = cfdart.configs.euler_discs(64, 4, (0.05, 0.2), (0, 1, 0, 1), seed=5)
    b = cfdart.Batch(2)
    s = b.add(d, h, max_steps=100000, T=0.02, cfl=0.8)
    b.step(200)
    t, n, act = b.status(s)
    ora = oracle.Sim(d, h)
    k, _ = ora.run(cfdart.ALTERNATING, 100000, T=0.02, cfl=0.8)
    assert not act and n == k and t == ora.t == 0.02
    assert bitwise_equal(interior(b.download(s)), interior(ora.grid))


def test_batch_admission_budget(gpu):
    d, h = cfdart.configs.euler_discs(64, 2, (0.05, 0.2), (0, 0, 0, 0), seed=1)
    one = 2 * (2 << 20)  # two state buffers of 64x64x4 doubles, rounded to 2 MiB each
    b = cfdart.Batch(8, mem_budget_bytes=3 * one)
    slots = [b.try_add(d, h, max_steps=5) for _ in range(4)]
    assert slots[:3] == [0, 1, 2] and slots[3] is None
    b.remove(1)
    assert b.try_add(d, h, max_steps=5) == 1
    assert b.info()[0] == 3


def test_unphysical_allaire_stops_alone_and_batched(gpu):
    """An Allaire bubble state run at CFL 2.5 (beyond the scheme's stability limit) fails the
    physical-state test during one of its first steps; the oracle counts the failed cells of that
    step. The reference's assert would abort there: the lone run ends after that step and returns
    CFDART_E_UNPHYS (synchronously, and at the next sync of an asynchronous run), the same grid in
    a batch stops at the same step with status flag 2, and every copy holds the oracle's state
    after that step, bitwise."""
    d, h = cfdart.configs.allaire_bubbles(48, 3, seed=9)
    ora = oracle.Sim(d, h)
    k = 0
    while ora.counters()[3] == 0:
        ora.run(cfdart.ALTERNATING, 1, cfl=2.5)
        k += 1
        assert k < 30
    want = interior(ora.grid)

    g = cfdart.Grid(d)
    g.upload(h)
    with pytest.raises(cfdart.CfdartError) as e:
        g.run(cfdart.ALTERNATING, 30, cfl=2.5)
    assert e.value.code == -4
    assert bitwise_equal(interior(g.download()), want)
    assert g.stats()[3] == ora.counters()[3]
    lone_sum = g.checksum()
    g.close()

    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 30, cfl=2.5, sync=False)
    with pytest.raises(cfdart.CfdartError) as e:
        g.sync()
    assert e.value.code == -4
    assert bitwise_equal(interior(g.download()), want)
    g.close()

    b = cfdart.Batch(4)
    s = b.add(d, h, splitting=cfdart.ALTERNATING, max_steps=30, cfl=2.5)
    b.step(30)
    b.sync()
    steps, flags = b.poll(4)
    assert steps[s] == k and flags[s] == 2
    assert bitwise_equal(interior(b.download(s)), want)
    assert b.checksum(s) == lone_sum
    b.close()


def test_batch_refuses_what_create_refuses(gpu):
    """Batch admission applies cfdart_create's descriptor checks (gamma > 1, wall codes,
    limiter and wave-speed options, PVRS on Euler, the z clip range, the grid size limit)."""
    de, he = cfdart.configs.euler_discs(32, 2, (0.05, 0.1), (0, 0, 0, 0), seed=1)
    da, ha = cfdart.configs.allaire_bubbles(32, 2, seed=1)
    bad = [(de, he, "gamma1", 1.0), (de, he, "limiter", 7), (de, he, "wave_speeds", 1), (da, ha, "wave_speeds", 3),
           (da, ha, "gamma2", 0.5), (da, ha, "z_lo", 1.0), (da, ha, "z_hi", 1.5)]
    b = cfdart.Batch(4)
    for d0, h, field, value in bad:
        d = type(d0).from_buffer_copy(d0)
        setattr(d, field, value)
        with pytest.raises(cfdart.CfdartError) as e:
            cfdart.Grid(d)
        assert e.value.code == -1, field
        with pytest.raises(cfdart.CfdartError) as e:
            b.add(d, h, max_steps=2)
        assert e.value.code == -1, field
    d = type(de).from_buffer_copy(de)
    d.bc[2] = 2
    with pytest.raises(cfdart.CfdartError) as e:
        b.add(d, he, max_steps=2)
    assert e.value.code == -1
    assert b.info()[0] == 0
    b.close()


def test_async_admission_removal_and_checksum(gpu):
    """Generated admissions, removals and checksums queued without waiting: a grid removed while
    steps are still queued keeps its state for them, a generated grid admitted right after into its
    memory starts from its own initial state, and a checksum queued between steps reports the
    state at that point of the stream (read after the next poll), whatever is queued after it."""
    import ctypes as C
    L = cfdart.lib()
    d, p = cfdart.configs.euler_discs_params(96, 4, (0.05, 0.15), seed=11)
    d2, p2 = cfdart.configs.euler_discs_params(96, 4, (0.05, 0.15), seed=12)
    b = cfdart.Batch(4)
    s0 = b.add_generated(d, p, max_steps=30, cfl=0.8)
    b.step(6)
    buf = L.cfdart_host_alloc(8)
    assert buf
    try:
        cfdart._lib._check(L.cfdart_batch_checksum_async(b.h, s0, C.c_void_p(buf)))
        b.step(4)                       # queued after the checksum
        b.remove(s0)                    # queued after those steps
        s1 = b.add_generated(d2, p2, max_steps=5, cfl=0.8)  # may reuse s0's slot and memory
        b.step(5)
        steps, flags = b.poll(4)
        print(f"case async batch ops: slots {s0} then {s1}", flush=True)
        got = C.c_ulonglong.from_address(buf).value
    finally:
        L.cfdart_host_free(C.c_void_p(buf))
    h = cfdart.random_discs(d, p)
    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 6, cfl=0.8)
    assert got == g.checksum()
    g.close()
    h2 = cfdart.random_discs(d2, p2)
    ref2, _, n2 = alone(d2, h2, 0.8, 5)
    assert steps[s1] == n2 == 5 and flags[s1] == 0
    assert bitwise_equal(interior(b.download(s1)), interior(ref2))
    b.close()


def test_step_shapes_follow_each_grid_parity(gpu):
    """A batch launches only the sweeps with work when every grid sweeps the same axis first (its
    step shape follows the grids' admission steps): a grid admitted at an odd batch step, left
    alone once the grids admitted at even steps are removed, still ends bitwise where it ends alone;
    so does a grid admitted later at an even step beside it (both shapes in one batch)."""
    cases = mixed_cases(4, 4)
    b = cfdart.Batch(4)
    (d0, h0, c0, _), (d1, h1, c1, _), (d2, h2, c2, _) = cases[:3]
    s0 = b.add(d0, h0, max_steps=3, cfl=c0)       # admitted at step 0
    b.step(3)
    s1 = b.add(d1, h1, max_steps=17, cfl=c1)      # admitted at step 3: sweeps y first on even steps
    b.step(2)
    b.remove(s0)                                  # only the odd-admitted grid is left
    b.step(6)
    s2 = b.add(d2, h2, max_steps=9, cfl=c2)       # admitted at step 11: x first on odd steps, like s1
    b.step(5)
    s3 = b.add(*cases[3][:2], max_steps=5, cfl=cases[3][2])  # admitted at step 16: the other parity
    b.step(20)
    b.sync()
    for (d, h, cfl, steps), s in (((d1, h1, c1, 17), s1), ((d2, h2, c2, 9), s2), (tuple(cases[3][:3]) + (5,), s3)):
        ref, t, n = alone(d, h, cfl, steps)
        tb, nb, act = b.status(s)
        print(f"case shapes slot={s} model={d.model} {d.nx}x{d.ny} steps={steps}", flush=True)
        assert nb == n == steps and tb == t and not act
        assert bitwise_equal(interior(b.download(s)), interior(ref))
    b.close()


def test_naive_and_alternating_grids_together(gpu):
    """Naive grids sweep x first on every step, alternating ones on every other: a batch holding
    both launches only x-first steps on even steps and both axes on odd ones; every grid still ends
    bitwise where it ends alone."""
    cases = mixed_cases(5, 4)
    b = cfdart.Batch(4)
    splits = [cfdart.ALTERNATING, cfdart.NAIVE, cfdart.ALTERNATING, cfdart.NAIVE]
    slots = [b.add(d, h, splitting=sp, max_steps=steps, cfl=cfl) for (d, h, cfl, steps), sp in zip(cases, splits)]
    b.step(30)
    b.sync()
    for (d, h, cfl, steps), sp, s in zip(cases, splits, slots):
        ref, t, n = alone(d, h, cfl, steps, sp)
        tb, nb, act = b.status(s)
        print(f"case naive+alternating slot={s} splitting={sp} {d.nx}x{d.ny} steps={steps}", flush=True)
        assert nb == n == steps and tb == t and not act
        assert bitwise_equal(interior(b.download(s)), interior(ref))
    b.close()


def test_strang_grids_among_the_others(gpu):
    """Strang grids (x(dt/2), y(dt), x(dt/2): three sweeps, so their state changes buffer at every
    step) in a batch with naive and alternating grids, one alternating grid admitted at an odd
    batch step: every grid ends bitwise where it ends alone, and a frame and a checksum queued
    after an odd and after an even number of a Strang grid's steps hold its state at that point."""
    import ctypes as C
    L = cfdart.lib()
    cases = mixed_cases(6, 5)
    splits = [cfdart.STRANG, cfdart.NAIVE, cfdart.STRANG, cfdart.ALTERNATING, cfdart.STRANG]
    runs = [(d, h, cfl, max(steps, 6)) for d, h, cfl, steps in cases]
    b = cfdart.Batch(8)
    slots = [b.add(d, h, splitting=sp, max_steps=n, cfl=cfl) for (d, h, cfl, n), sp in zip(runs[:3], splits[:3])]
    d0, h0, c0, _ = runs[0]
    nb = 8 * d0.nx * d0.ny * (6 if d0.model == 1 else 4)
    host = L.cfdart_host_alloc(2 * nb + 16)
    assert host
    try:
        b.step(3)                                   # Strang grid 0 has taken 3 steps: state in its second buffer
        b.frame(slots[0], host)
        cfdart._lib._check(L.cfdart_batch_checksum_async(b.h, slots[0], C.c_void_p(host + 2 * nb)))
        for (d, h, cfl, n), sp in zip(runs[3:], splits[3:]):   # admitted at batch step 3
            slots.append(b.add(d, h, splitting=sp, max_steps=n, cfl=cfl))
        b.step(1)                                   # 4 steps: back in its first buffer
        assert b.frames([slots[0]], [host + nb]) == [0]
        cfdart._lib._check(L.cfdart_batch_checksum_async(b.h, slots[0], C.c_void_p(host + 2 * nb + 8)))
        b.step(40)
        b.frames_done()
        b.sync()
        nc = 6 if d0.model == 1 else 4
        frames = [np.ctypeslib.as_array((C.c_double * (nb // 8)).from_address(host + o)).reshape(nc, d0.ny, d0.nx).copy()
                  for o in (0, nb)]
        sums = [C.c_ulonglong.from_address(host + 2 * nb + 8 * i).value for i in range(2)]
    finally:
        L.cfdart_host_free(C.c_void_p(host))
    g = cfdart.Grid(d0)
    g.upload(h0)
    for k, (n, fr, cs) in enumerate(zip((3, 1), frames, sums)):
        g.run(cfdart.STRANG, n, cfl=c0)
        st = g.download()
        print(f"case strang mid-run slot={slots[0]} after {3 + k} steps", flush=True)
        assert cs == g.checksum()
        assert np.array_equal(fr.view(np.uint64), np.moveaxis(st[2:-2, 2:-2, :], 2, 0).view(np.uint64))
    g.close()
    for (d, h, cfl, n), sp, s in zip(runs, splits, slots):
        ref, t, nn = alone(d, h, cfl, n, sp)
        tb, nb_, act = b.status(s)
        print(f"case strang batch slot={s} splitting={sp} model={d.model} {d.nx}x{d.ny} steps={n}", flush=True)
        assert nb_ == nn == n and tb == t and not act
        assert bitwise_equal(interior(b.download(s)), interior(ref))
        assert b.checksum(s) == cfdart.state_checksum(ref)
    b.close()
