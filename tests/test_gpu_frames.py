"""GPU frame products: the device schlieren image against its definition evaluated on the host
from the downloaded state (bitwise: same IEEE operations), and the asynchronous state frame
against the downloaded state."""
import ctypes as C

import numpy as np
import pytest

import cfdart

pytestmark = pytest.mark.gpu


def host_schlieren(d, state):
    g = d.ng
    rho = state[..., 0] + (state[..., 1] if d.model == 1 else 0.0)
    rho_L, rho_R = rho[g:-g, g - 1:-g - 1], rho[g:-g, g + 1:(-g + 1) or None]
    rho_B, rho_T = rho[g - 1:-g - 1, g:-g], rho[g + 1:(-g + 1) or None, g:-g]
    gradx = (rho_R - rho_L) / (2.0 * d.dx)
    grady = (rho_T - rho_B) / (2.0 * d.dy)
    v = np.sqrt(gradx * gradx + grady * grady)
    return v / v.max(), v.max()


@pytest.mark.parametrize("model,bc", [(0, (0, 0, 0, 0)), (0, (1, 1, 0, 1)), (1, (0, 0, 0, 0)), (1, (1, 0, 1, 0))])
def test_schlieren_matches_definition(gpu, model, bc):
    if model == 0:
        d, h = cfdart.configs.euler_discs(150, 8, (0.03, 0.1), bc, seed=3, ny=110)
        cfl = 0.8
    else:
        d, h = cfdart.configs.allaire_bubbles(150, 6, seed=3, ny=110)
        for i in range(4):
            d.bc[i] = bc[i]
        cfl = 0.5
    print(f"case schlieren model={model} bc={bc}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 7, cfl=cfl)
    img, m = g.schlieren()
    ref, mref = host_schlieren(d, g.download())
    assert m == mref
    assert np.array_equal(img, ref)


def test_state_frame_matches_download(gpu):
    d, h = cfdart.configs.allaire_bubbles(90, 4, seed=8, ny=70)
    g = cfdart.Grid(d)
    g.upload(h)
    g.run(cfdart.ALTERNATING, 5, cfl=0.5)
    n = d.nx * d.ny * 6
    p = cfdart.lib().cfdart_host_alloc(8 * n)
    assert p
    try:
        assert cfdart.lib().cfdart_frame(g.h, C.c_void_p(p)) == 0
        g.sync()
        frame = np.ctypeslib.as_array((C.c_double * n).from_address(p)).reshape(6, d.ny, d.nx).copy()
    finally:
        cfdart.lib().cfdart_host_free(C.c_void_p(p))
    s = g.download()[2:-2, 2:-2]
    assert np.array_equal(frame, np.moveaxis(s, 2, 0))


@pytest.mark.parametrize("model", [0, 1])
def test_mass_history_matches_compute_dt(gpu, model):
    """The device mass history of a run equals, step by step, the masses compute_dt records (the
    library's own cfdart_compute_dt bitwise: same partials and order) and the oracle's row-order
    sums to round-off; with reflective walls all round the masses stay constant to round-off."""
    import math

    import oracle
    if model == 0:
        d, h = cfdart.configs.euler_discs(120, 6, (0.05, 0.15), (1, 1, 1, 1), seed=4, ny=90)
        cfl = 0.8
    else:
        d, h = cfdart.configs.allaire_bubbles(120, 5, seed=4, ny=90)
        cfl = 0.5
    print(f"case mass history model={model}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n, hist = g.run_mass(cfdart.ALTERNATING, 25, cfl=cfl)
    assert n == 25 and hist.shape == (25, 3)
    ora = oracle.Sim(d, h)
    step_g = cfdart.Grid(d)
    step_g.upload(h)
    for k in range(25):
        dt, (m1, m2) = ora.compute_dt(cfl=cfl)
        gdt, gm = step_g.compute_dt(k, ora.t, math.inf, cfl)
        assert hist[k, 0] == ora.t
        assert hist[k, 1] == gm[0] and (model == 0 or hist[k, 2] == gm[1])
        assert abs(hist[k, 1] - m1) <= 1e-13 * abs(m1)
        if model == 1:
            assert abs(hist[k, 2] - m2) <= 1e-13 * abs(m2)
        ora.advance(cfdart.ALTERNATING, dt)
        step_g.step(cfdart.ALTERNATING, k, gdt)
        ora.t += dt
        ora.n += 1
    if model == 0:  # closed box: mass conserved to round-off
        assert np.max(np.abs(hist[:, 1] - hist[0, 1])) <= 1e-13 * hist[0, 1]


def _frame_state(planes, d):
    """A host state (cfdart.h layout) whose interior is a frame's planes ([component][y][x])."""
    s = np.zeros((d.ny + 2 * d.ng, d.nx + 2 * d.ng, planes.shape[0]))
    s[d.ng:-d.ng, d.ng:-d.ng, :] = np.moveaxis(planes, 0, 2)
    return s


def test_batch_frames_match_state(gpu):
    """Batch frames queued between steps without waiting (staged in the device ring, copied to
    the host on the copy stream) hold each grid's state at their point of the stream: a frame's
    checksum equals the device checksum queued right after it. The small grids' frames wrap the
    128 MiB staging ring while earlier copies are still in flight; a large grid joining later
    regrows the ring. Frames are taken one per call and in groups (cfdart_batch_frames, one event
    between the streams per group). The last frame of each grid equals its downloaded state."""
    L = cfdart.lib()
    specs = [(0, 300, 300), (1, 257, 200), (0, 130, 130), (1, 64, 64)]
    b = cfdart.Batch(8)

    def make(model, nx, ny, seed):
        if model == 0:
            d, h = cfdart.configs.euler_discs(nx, 4, (0.05, 0.15), (0, 0, 0, 0), seed=seed, ny=ny)
            return d, h, 0.8
        d, h = cfdart.configs.allaire_bubbles(nx, 3, seed=seed, ny=ny)
        return d, h, 0.5

    def nbytes(d):
        return 8 * d.nx * d.ny * (6 if d.model == 1 else 4)

    grids = []
    for k, (model, nx, ny) in enumerate(specs):
        d, h, cfl = make(model, nx, ny, 40 + k)
        grids.append((b.add(d, h, max_steps=100, cfl=cfl), d))
    dl, hl, cl = make(1, 1190, 1190, 50)
    rounds1, rounds2 = 24, 3
    small = sum(nbytes(d) for _, d in grids)
    total = rounds1 * small + (rounds2 + 1) * (small + nbytes(dl))
    nframes = rounds1 * len(grids) + (rounds2 + 1) * (len(grids) + 1)
    assert rounds1 * small > 128 * 2**20  # the first phase wraps the staging ring
    host = L.cfdart_host_alloc(total)
    sums = L.cfdart_host_alloc(8 * nframes)
    assert host and sums
    taken = []  # (slot, desc, byte offset, frame index)
    try:
        off = 0

        def frame_all(group):
            # one frame per call, or the frames of every grid as one group (cfdart_batch_frames)
            nonlocal off
            first = len(taken)
            if group:
                ptrs = []
                for s, d in grids:
                    ptrs.append(host + off)
                    taken.append((s, d, off, len(taken)))
                    off += nbytes(d)
                assert b.frames([s for s, _ in grids], ptrs) == [0] * len(grids)
            for k, (s, d) in enumerate(grids):
                if not group:
                    b.frame(s, host + off)
                    taken.append((s, d, off, len(taken)))
                    off += nbytes(d)
                i = first + k
                cfdart._lib._check(L.cfdart_batch_checksum_async(b.h, s, C.c_void_p(sums + 8 * i)))

        for r in range(rounds1):
            frame_all(r % 2 == 1)
            b.step(1)
        grids.append((b.add(dl, hl, max_steps=100, cfl=cl), dl))
        for r in range(rounds2):
            frame_all(r % 2 == 0)  # the first group regrows the ring after staging the small frames
            b.step(1)
        frame_all(True)  # the final states, compared with downloads below
        print(f"case batch frames: {len(taken)} frames, {off / 2**20:.0f} MiB", flush=True)
        b.frames_done()
        b.sync()
        last = {}
        for s, d, o, i in taken:
            nc = 6 if d.model == 1 else 4
            planes = np.ctypeslib.as_array((C.c_double * (nc * d.ny * d.nx)).from_address(host + o))
            planes = planes.reshape(nc, d.ny, d.nx).copy()
            want = C.c_ulonglong.from_address(sums + 8 * i).value
            assert cfdart.state_checksum(_frame_state(planes, d)) == want, (s, i)
            last[s] = planes
        for s, d in grids:
            st = b.download(s)
            assert np.array_equal(last[s].view(np.uint64), np.moveaxis(st[d.ng:-d.ng, d.ng:-d.ng, :], 2, 0).view(np.uint64))
    finally:
        L.cfdart_host_free(C.c_void_p(host))
        L.cfdart_host_free(C.c_void_p(sums))
    b.close()


def test_batch_frames_skip_what_the_budget_cannot_stage(gpu):
    """A frame whose staging would grow the device ring past the batch's memory budget is skipped
    in a group (status CFDART_E_NOMEM) while the group's other frames are taken; alone it is
    refused with that code."""
    L = cfdart.lib()
    ds, hs = cfdart.configs.euler_discs(200, 4, (0.05, 0.15), (0, 0, 0, 0), seed=61)
    dl, hl = cfdart.configs.allaire_bubbles(1190, 3, seed=62)      # a frame of 8.5 M doubles: the ring must grow
    def state(d):  # a grid's two state buffers as the batch counts them (256-byte rows, 2 MiB granules)
        one = 8 * ((d.nx + 31) // 32 * 32) * d.ny * (6 if d.model == 1 else 4)
        return 2 * (-(-one // (2 << 20)) * (2 << 20))
    budget = 2 * state(ds) + state(dl) + 8 * 2**24 + (4 << 20)       # states, the first 128 MiB ring, some slack
    b = cfdart.Batch(4, mem_budget_bytes=budget)
    s0 = b.add(ds, hs, max_steps=10, cfl=0.8)
    s1 = b.add(dl, hl, max_steps=10, cfl=0.5)
    s2 = b.add(ds, hs, max_steps=10, cfl=0.8)
    b.step(2)
    nb = 8 * ds.nx * ds.ny * 4
    host = L.cfdart_host_alloc(3 * nb)
    assert host
    try:
        b.frame(s0, host + 2 * nb)                                     # a small frame: the 128 MiB ring
        st = b.frames([s0, s1, s2], [host, host, host + nb])
        print(f"case frames under budget: status {st}", flush=True)
        assert st == [0, -2, 0]
        with pytest.raises(cfdart.CfdartError) as e:
            b.frame(s1, host)
        assert e.value.code == -2
        b.frames_done()
        ref = np.moveaxis(b.download(s0)[2:-2, 2:-2, :], 2, 0)
        for o in (0, nb, 2 * nb):
            fr = np.ctypeslib.as_array((C.c_double * (nb // 8)).from_address(host + o)).reshape(4, ds.ny, ds.nx)
            assert np.array_equal(fr.view(np.uint64), ref.view(np.uint64))
    finally:
        L.cfdart_host_free(C.c_void_p(host))
    b.close()
