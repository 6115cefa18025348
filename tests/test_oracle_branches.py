"""Pins of the oracle's rarely taken branches and of its splitting composition, against what the
reference's semantics fix independently of the oracle's arithmetic:

* the Euler positivity repair (E/sourcecode/euler_rowsweep.cpp:47-82, column variant
  E/sourcecode/euler_colsweep.cpp:78-113): which state each recovered quantity comes from, and
  that both thresholds are inclusive (<= 1e-12), checked at the thresholds from both sides;
* the MUSCL-Hancock face's fallbacks (E/headers/flux_solver_MUSCLhancock.hpp:178-223,
  A/headers/flux_solver_MUSCLhancock.hpp:324-357): M3 zeroes both slopes, so the face must equal
  the HLLC flux of the two cells' own primitive states re-converted; M5 must equal the HLLC flux
  of the two raw stencil cells s1, s2 (unsnapped); Allaire's snapping overload decides M5 and
  the snapped state is what HLLC receives;
* the splitting schemes (E/headers/splitting_scheme.hpp:37-122): composed against single
  sweeps, against a state on which the y sweep is the identity (so the dt fractions of the x
  sweeps show), and against the transposed problem (so the sweep order shows).

Planted mistakes these tests catch (tests/test_oracle_mutants.py builds each one into the oracle
and runs the test named beside it):
  repair  `rho <= 1e-12` -> `<`                      test_repair_thresholds_inclusive
  repair  `e <= 1e-12` -> `<`                        test_repair_thresholds_inclusive
  repair  e_old from the new state                   test_repair_sources
  repair  velocities from the old state              test_repair_sources
  M3      only the left slope zeroed                 test_slope_fallback_allaire
  M5      fallback to s0, s3 instead of s1, s2       test_evolve_fallback_euler
  M5      const (non-snapping) physical test         test_evolve_snapping_allaire
  Strang  X(dt) Y(dt) X(dt)                          test_splitting_y_identity_state
  alternating  odd steps in the even steps' order    test_alternating_odd_is_transposed_naive
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle

L = oracle.lib()

EUL = oracle.make_desc(model=0, nx=8, ny=8, dx=0.1, dy=0.1, gamma1=1.4, pinf1=0.0)
SG = oracle.make_desc(model=0, nx=8, ny=8, dx=0.1, dy=0.1, gamma1=4.4, pinf1=6e8)
ALL = oracle.make_desc(model=1, nx=8, ny=8, dx=0.1, dy=0.1, gamma1=4.4, pinf1=6e8, gamma2=1.4, pinf2=0.0)


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def vec(x):
    return np.ascontiguousarray(np.array(x, dtype=np.float64))


def c2p(d, U):
    U = vec(U)
    W = np.zeros_like(U)
    L.orc_c2p(C.byref(d), dp(U), dp(W))
    return W


def p2c(d, W):
    W = vec(W)
    U = np.zeros_like(W)
    L.orc_p2c(C.byref(d), dp(W), dp(U))
    return U


def is_phys_snap(d, U):
    U = vec(U).copy()
    ok = L.orc_is_physical_snap(C.byref(d), dp(U))
    return bool(ok), U


def hllc(d, UL, UR):
    UL, UR = vec(UL), vec(UR)
    F = np.zeros_like(UL)
    us = C.c_double()
    L.orc_hllc(C.byref(d), dp(UL), dp(UR), dp(F), C.byref(us))
    return F, us.value


def face(d, s0, s1, s2, s3, dt, dx):
    s = [vec(x) for x in (s0, s1, s2, s3)]
    F = np.zeros_like(s[0])
    us, zs = C.c_double(), C.c_double()
    cnt = (C.c_longlong * 4)()
    L.orc_face(C.byref(d), dp(s[0]), dp(s[1]), dp(s[2]), dp(s[3]), dt, dx, dp(F), C.byref(us), C.byref(zs), cnt)
    return F, us.value, zs.value, list(cnt)


def repair(d, old, fut):
    old, fut = vec(old), vec(fut).copy()
    r = L.orc_euler_repair(C.byref(d), dp(old), dp(fut))
    return r, fut


def e_of(d, U):
    """Specific internal energy of a conserved Euler state (closed form, no oracle code)."""
    rho, mu, mv, E = U
    return E / rho - 0.5 * ((mu / rho) ** 2 + (mv / rho) ** 2)


def e_close(d, U, e_want):
    """e of the repaired state against its source value. The repair goes through the pressure,
    p = (g-1) rho e - g pinf, and back, so with pinf > 0 the recovered e carries the rounding of
    g pinf relative to (g-1) rho: a few ulp of g pinf / ((g-1) rho)."""
    tol = 1e-13 * abs(e_want) + 8 * 2.3e-16 * d.gamma1 * d.pinf1 / ((d.gamma1 - 1.0) * U[0])
    return abs(e_of(d, U) - e_want) <= tol


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


# ---------------------------------------------------------------- Euler repair

OLD_W = [0.8, 0.3, -0.2, 2.0]


@pytest.mark.parametrize("d", [EUL, SG], ids=["ideal", "stiffened"])
def test_repair_leaves_physical_states(d):
    old = p2c(d, OLD_W)
    for W in ([1.0, 0.0, 0.0, 1.0], [1e-12, 5.0, -3.0, 1e-3], [2.0, -40.0, 7.0, 1e6]):
        fut = p2c(d, W)
        r, out = repair(d, old, fut)
        assert r == 0 and np.array_equal(bits(out), bits(fut))


@pytest.mark.parametrize("d", [EUL, SG], ids=["ideal", "stiffened"])
def test_repair_thresholds_inclusive(d):
    """rho <= 1e-12 and e <= 1e-12 both take the old value; just above either, the new value is
    kept. (The reference's comparisons are <=; a strict < fails the first and third cases.)"""
    old = p2c(d, OLD_W)
    e_old = e_of(d, old)
    gp = d.gamma1 * d.pinf1
    # rho' = 1e-12 exactly, e' = -1: rho and e from the old state
    r, out = repair(d, old, [1e-12, 0.0, 0.0, -1e-12])
    assert r == 1 and out[0] == old[0]
    assert e_close(d, out, e_old)
    # rho' just above 1e-12, e' = -1: rho kept, e from the old state
    rho = math.nextafter(1e-12, 1.0)
    r, out = repair(d, old, [rho, 0.0, 0.0, -rho])
    assert r == 1 and out[0] == rho
    assert e_close(d, out, e_old)
    # rho' < 0 (unphysical) with e' = 1e-12 exactly: e from the old state
    r, out = repair(d, old, [-1.0, 0.0, 0.0, -1e-12])
    assert r == 1 and out[0] == old[0]
    assert e_close(d, out, e_old)
    # rho' < 0 with e' just above 1e-12: e kept (p from rho_old and e': (g-1) rho e - g pinf)
    e_new = math.nextafter(1e-12, 1.0)
    r, out = repair(d, old, [-1.0, 0.0, 0.0, -e_new])
    assert r == 1 and out[0] == old[0]
    p = c2p(d, out)[3]
    assert math.isclose(p + gp, (d.gamma1 - 1.0) * old[0] * e_new, rel_tol=1e-9, abs_tol=1e-12 * max(gp, 1.0))


@pytest.mark.parametrize("d", [EUL, SG], ids=["ideal", "stiffened"])
def test_repair_sources(d):
    """The repaired state keeps the new velocities (u = m'/rho', with the new rho) and takes the
    old state's specific internal energy; the old velocities enter only through e_old."""
    old = p2c(d, OLD_W)
    e_old = e_of(d, old)
    rho, u, v = 0.5, 3.0, -1.0
    fut = [rho, rho * u, rho * v, 0.5 * rho * (u * u + v * v) - rho * 1.0]  # e' = -1
    r, out = repair(d, old, fut)
    assert r == 1
    assert out[0] == rho
    assert math.isclose(out[1] / out[0], u, rel_tol=1e-15) and math.isclose(out[2] / out[0], v, rel_tol=1e-15)
    assert e_close(d, out, e_old)
    # density below the threshold: rho_old with the velocities of the new state (m'/rho')
    fut = [1e-13, 1e-13 * u, 1e-13 * v, -1.0]
    r, out = repair(d, old, fut)
    assert r == 1 and out[0] == old[0]
    assert math.isclose(out[1] / out[0], u, rel_tol=1e-13) and math.isclose(out[2] / out[0], v, rel_tol=1e-13)
    assert e_close(d, out, e_old)


def test_repair_old_state_rotation_invariant():
    """The column sweep passes the unrotated old state (E/sourcecode/euler_colsweep.cpp:91-94);
    the repair reads only its density and specific internal energy, which the rotation
    (m_x, m_y) -> (m_y, -m_x) leaves bitwise unchanged, so either frame gives the same bits."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        W = [10 ** rng.uniform(-2, 2), rng.uniform(-9, 9), rng.uniform(-9, 9), 10 ** rng.uniform(-2, 3)]
        old = p2c(EUL, W)
        rot = old.copy()
        rot[1], rot[2] = old[2], -old[1]
        fut = [rng.uniform(-1, 1e-12), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)]
        _, a = repair(EUL, old, fut)
        _, b = repair(EUL, rot, fut)
        assert np.array_equal(bits(a), bits(b))


def test_repair_in_a_sweep():
    """A sweep with a time step far above the CFL limit drives cells unphysical: every interior
    cell equals the conservative update from the face fluxes followed by the repair against the
    cell's old state, and the repair counter counts exactly the repaired cells."""
    n = 24
    d = oracle.make_desc(model=0, nx=n, ny=3, dx=1.0 / n, dy=1.0 / n, gamma1=1.4, limiter=0)
    nc = 4
    g = np.zeros((3 + 4, n + 4, nc))
    for j in range(n + 4):
        x = (j - 2 + 0.5) / n
        W = [1.0, -3.0, 0.0, 0.4] if x < 0.5 else [0.05, 3.0, 0.0, 0.02]
        g[:, j] = p2c(d, W)
    dt = 1.5 / n / 4.0
    s = oracle.Sim(d, g)
    s.fill_ghosts()
    old = s.grid.copy()
    s.sweep(0, dt)
    got = s.grid
    nrep = 0
    for i in range(2, 5):
        row = old[i]
        flux = {}
        for j in range(2, n + 3):
            flux[j], _, _, _ = face(d, row[j - 2], row[j - 1], row[j], row[j + 1], dt, d.dx)
        for j in range(2, n + 2):
            raw = row[j] + (dt / d.dx) * (flux[j] - flux[j + 1])
            r, want = repair(d, row[j], raw)
            nrep += r
            assert np.array_equal(bits(got[i, j]), bits(want)), (i, j)
    assert nrep > 0 and s.counters()[2] == nrep


# ---------------------------------------------------------------- face fallbacks

def test_slope_fallback_allaire():
    """Negative pressures (tension) are physical in the stiffened mixture while p + k(z) > 0, and
    the sound speed stays real while p > -k(z) / (1 + 1/eps(z)). Four such cells (z from 0.02 to
    0.89, tensions up to 6.3e7 Pa) have a minmod extrapolation to the face that fails the
    physical-state test: the mixture k(z) is not linear in z, so the linear reconstruction of p
    and z leaves the set in which both endpoints lie. M3 must then zero both slopes, and the face
    is the HLLC flux of the two cells' own primitive states re-converted (the half step of a zero
    slope returns W exactly). The velocity slopes are non-zero on both sides, so zeroing only one
    slope changes the result."""
    d = ALL
    eps, k = C.c_double(), C.c_double()

    def mix(z):
        L.orc_mixture_eos(C.byref(d), z, C.byref(eps), C.byref(k))
        return k.value, eps.value

    Ws = [[840.0, 0.16, 0.0, 0.0, -1.188e7, 0.84], [250.0, 0.75, 0.5, 0.0, -6.3e7, 0.25],
          [20.0, 0.98, 1.0, 0.0, -3.1e5, 0.02], [890.0, 0.11, 1.5, 0.0, -4.5e6, 0.89]]
    for w in Ws:
        kz, ez = mix(w[5])
        assert w[4] + kz > 0 and w[4] > -kz / (1.0 + 1.0 / ez)  # physical, real sound speed
    s = [p2c(d, w) for w in Ws]
    W = [c2p(d, x) for x in s]
    delL = [L.orc_limit(0, W[1][j] - W[0][j], W[2][j] - W[1][j]) for j in range(6)]
    delR = [L.orc_limit(0, W[2][j] - W[1][j], W[3][j] - W[2][j]) for j in range(6)]
    assert delL[2] != 0.0 and delR[2] != 0.0
    okL, _ = is_phys_snap(d, p2c(d, W[1] + 0.5 * np.array(delL)))
    okR, _ = is_phys_snap(d, p2c(d, W[2] - 0.5 * np.array(delR)))
    assert not (okL and okR)  # an extrapolation leaves the physical set
    dt, dx = 1e-6, 0.1
    F, us, zs, cnt = face(d, *s, dt, dx)
    assert cnt[0] == 1 and cnt[1] == 0
    UL, UR = p2c(d, W[1]), p2c(d, W[2])
    Fw, usw = hllc(d, UL, UR)
    assert np.all(np.isfinite(Fw))
    assert np.array_equal(bits(F), bits(Fw)) and us == usw
    assert zs == (UL[5] if math.copysign(1.0, us) > 0 else UR[5])


def test_evolve_fallback_euler():
    """Diverging velocities with dt/dx = 4: the half step drives the evolved densities negative
    (rho - s rho du = 1 - 2), the extrapolations themselves stay physical, and M5 must give the
    HLLC flux of the two raw stencil cells s1, s2."""
    d = EUL
    Ws = [[1.0, -2.0, 0.0, 1.0], [1.0, -1.0, 0.3, 1.0], [1.0, 1.0, -0.2, 1.0], [1.0, 2.0, 0.0, 1.0]]
    s = [p2c(d, w) for w in Ws]
    F, us, _, cnt = face(d, *s, 0.4, 0.1)
    assert cnt[0] == 0 and cnt[1] == 1
    Fw, usw = hllc(d, s[1], s[2])
    assert np.array_equal(bits(F), bits(Fw)) and us == usw
    Fo, _ = hllc(d, s[0], s[3])
    assert not np.array_equal(bits(F), bits(Fo))


def test_evolve_fallback_allaire_raw_unsnapped():
    """Allaire M5 failure (the partial density of fluid 2 goes to -1 in the half step): the face is
    the HLLC flux of the raw cells s1, s2 as stored -- their 5e-13 partial density of fluid 1 is
    not snapped to 0."""
    d = ALL
    Ws = [[5e-13, 1.0, u, 0.0, 1e5, 0.5] for u in (-2.0, -1.0, 1.0, 2.0)]
    s = [p2c(d, w) for w in Ws]
    F, us, zs, cnt = face(d, *s, 0.4, 0.1)
    assert cnt[0] == 0 and cnt[1] == 1
    Fw, usw = hllc(d, s[1], s[2])
    assert np.array_equal(bits(F), bits(Fw)) and us == usw
    snapped = [x.copy() for x in s]
    for x in snapped:
        x[0] = 0.0
    Fs, _ = hllc(d, snapped[1], snapped[2])
    assert not np.array_equal(bits(F), bits(Fs))


def test_evolve_snapping_allaire():
    """The half step takes the partial density of fluid 1 from 1e-13 to -1.25e-13 on both sides;
    the snapping overload sets it to 0, the state passes, and HLLC receives the snapped evolved
    states (no fallback). Only component 0 has a slope here, so each evolved state is W with that
    component replaced: T0 = (W0 + 0.5 del0) - (s u) del0 (row 0 of the Jacobian is u, 0, W0, 0, 0,
    0 and every other slope is zero)."""
    d = ALL
    a = [0.5e-13, 1.0e-13, 1.5e-13, 2.0e-13]
    Ws = [[ai, 1.0, 1.0, 0.0, 1e5, 1e-6] for ai in a]
    s = [p2c(d, w) for w in Ws]
    dt, dx = 1.0, 0.1
    sc = 0.5 * (dt / dx)
    W0, W1, W2, W3 = (c2p(d, x) for x in s)
    delL = L.orc_limit(0, W1[0] - W0[0], W2[0] - W1[0])
    delR = L.orc_limit(0, W2[0] - W1[0], W3[0] - W2[0])
    TL, TR = W1.copy(), W2.copy()
    TL[0] = (W1[0] + 0.5 * delL) - (sc * W1[2]) * delL
    TR[0] = (W2[0] - 0.5 * delR) - (sc * W2[2]) * delR
    assert TL[0] < 0 and TR[0] < 0 and abs(TL[0]) < 1e-12 and abs(TR[0]) < 1e-12
    okL, UL = is_phys_snap(d, p2c(d, TL))
    okR, UR = is_phys_snap(d, p2c(d, TR))
    assert okL and okR and UL[0] == 0.0 and UR[0] == 0.0
    F, us, zs, cnt = face(d, *s, dt, dx)
    assert cnt == [0, 0, 0, 0]
    Fw, usw = hllc(d, UL, UR)
    assert np.array_equal(bits(F), bits(Fw)) and us == usw
    assert zs == (UL[5] if math.copysign(1.0, us) > 0 else UR[5])


# ---------------------------------------------------------------- splitting

def disc_state(n, model, bc, seed):
    import cfdart
    if model == 0:
        d, h = cfdart.configs.euler_discs(n, 4, (0.08, 0.2), bc, seed=seed)
    else:
        d, h = cfdart.configs.allaire_bubbles(n, 3, seed=seed)
        for i in range(4):
            d.bc[i] = bc[i]
    return oracle.desc_from(d), np.ascontiguousarray(h)


def transpose_state(h, model):
    t = np.ascontiguousarray(np.transpose(h, (1, 0, 2))).copy()
    mx = 2 if model else 1
    t[..., [mx, mx + 1]] = t[..., [mx + 1, mx]]
    return t


def transpose_desc(d):
    t = oracle.desc_from(d)
    t.nx, t.ny, t.dx, t.dy = d.ny, d.nx, d.dy, d.dx
    t.bc[0], t.bc[1], t.bc[2], t.bc[3] = d.bc[2], d.bc[3], d.bc[0], d.bc[1]
    return t


@pytest.mark.parametrize("model", [0, 1])
def test_splitting_composed_from_sweeps(model):
    """orc_advance equals its sequence of (ghost fill, sweep, swap) calls with the reference's
    axes and dt fractions: naive X(dt) Y(dt); Strang X(dt/2) Y(dt) X(dt/2); alternating X Y on
    even steps and Y X on odd ones (bitwise, counters equal)."""
    d, h = disc_state(32, model, (1, 0, 0, 1), 3)
    dt = 2e-4 if model == 0 else 2e-7
    plans = {(0, 0): [(0, dt), (1, dt)], (1, 0): [(0, 0.5 * dt), (1, dt), (0, 0.5 * dt)],
             (2, 0): [(0, dt), (1, dt)], (2, 1): [(1, dt), (0, dt)]}
    for (split, n), plan in plans.items():
        a = oracle.Sim(d, h)
        a.advance(split, dt, n=n)
        b = oracle.Sim(d, h)
        for axis, dts in plan:
            b.sweep(axis, dts)
        assert np.array_equal(bits(a.grid[2:-2, 2:-2]), bits(b.grid[2:-2, 2:-2])), (split, n)
        assert a.counters() == b.counters()


@pytest.mark.parametrize("model", [0, 1])
def test_splitting_y_identity_state(model):
    """On a state that is uniform along y with v = 0 the y sweep is the identity (bitwise), so a
    step shows its x sweeps: naive and alternating give X(dt), Strang gives X(dt/2) twice, which
    differs from X(dt)."""
    n = 40
    if model == 0:
        d = oracle.make_desc(model=0, nx=n, ny=6, dx=1.0 / n, dy=1.0 / n, gamma1=1.4, limiter=1)
        WL, WR, dt = [1.0, 0.2, 0.0, 1.0], [0.125, -0.1, 0.0, 0.1], 1e-3
    else:
        d = oracle.make_desc(model=1, nx=n, ny=6, dx=1.0 / n, dy=1.0 / n, gamma1=4.4, pinf1=6e8, gamma2=1.4,
                             z_lo=1e-6, z_hi=1 - 1e-6, wave_speeds=1)
        WL = [1000 * (1 - 1e-6), 1e-6, 10.0, 0.0, 1e9, 1 - 1e-6]
        WR = [1e-3, 50.0, -5.0, 0.0, 1e5, 1e-6]  # non-zero u: the rotation would sign a zero
        dt = 2e-6
    nc = len(WL)
    h = np.zeros((6 + 4, n + 4, nc))
    for j in range(n + 4):
        h[:, j] = p2c(d, WL if j < n // 2 + 2 else WR)
    y = oracle.Sim(d, h)
    y.sweep(1, dt)
    assert np.array_equal(bits(y.grid[2:-2, 2:-2]), bits(h[2:-2, 2:-2]))

    def run(split, n_step=0):
        s = oracle.Sim(d, h)
        s.advance(split, dt, n=n_step)
        return s.grid[2:-2, 2:-2].copy()

    x1 = oracle.Sim(d, h)
    x1.sweep(0, dt)
    x2 = oracle.Sim(d, h)
    x2.sweep(0, 0.5 * dt)
    x2.sweep(0, 0.5 * dt)
    X, XX = x1.grid[2:-2, 2:-2], x2.grid[2:-2, 2:-2]
    assert not np.array_equal(bits(X), bits(XX))
    assert np.array_equal(bits(run(0)), bits(X))
    assert np.array_equal(bits(run(2, 0)), bits(X))
    assert np.array_equal(bits(run(2, 1)), bits(X))
    assert np.array_equal(bits(run(1)), bits(XX))


@pytest.mark.parametrize("model", [0, 1])
def test_alternating_odd_is_transposed_naive(model):
    """With square cells, an odd alternating step (Y then X) of a state equals the transpose of a
    naive step (X then Y) of the transposed state, walls transposed with it; the even alternating
    step equals the naive step itself. The two orders differ on a genuinely 2D state. Values are
    compared exactly with +0 == -0: the column sweep's rotation (m_x, m_y) -> (m_y, -m_x) gives a
    cell at rest a negative zero where the row sweep keeps a positive one."""
    d, h = disc_state(36, model, (1, 0, 0, 1), 11)
    dt = 3e-4 if model == 0 else 3e-7
    td, th = transpose_desc(d), transpose_state(h, model)

    def step(desc, state, split, n):
        s = oracle.Sim(desc, state)
        s.advance(split, dt, n=n)
        return s.grid[2:-2, 2:-2].copy()

    odd = step(d, h, 2, 1)
    even = step(d, h, 2, 0)
    naive_t = transpose_state(np.pad(step(td, th, 0, 0), ((2, 2), (2, 2), (0, 0))), model)[2:-2, 2:-2]
    assert np.array_equal(odd, naive_t)
    assert np.array_equal(bits(even), bits(step(d, h, 0, 0)))
    assert not np.array_equal(odd, even)
