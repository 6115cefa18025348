"""Named workloads: the benchmark's configurations as grid descriptors + initial conditions.

euler_200      Euler, gamma 1.4, 200x200 on [0,1]^2, ambient (1,0,0,1), 8 seeded discs
               (centres U[0.1,0.9]^2, r U[0.03,0.10], p U[5,20], rho U[0.5,2]), transmissive,
               van Leer, HLLC, alternating splitting, CFL 0.8, run to t = 0.2.
euler_2048     As above at 2048x2048 with 64 discs of r U[0.005,0.03]; reflective left/right,
               transmissive bottom/top; 100 steps at CFL 0.8.
euler_slab     One 2048x256 y-slab of euler_2048 (one GPU's share of eight).
euler_2048_godunov  euler_2048 with the first-order Godunov flux.
allaire_1024   Allaire water (gamma 4.4, pinf 6e8, rho 1000) with 16 air bubbles (gamma 1.4,
               rho 1), p 1e5, radii U[0.02,0.06], centres U[0.15,0.95]x[0.05,0.95]; a planar
               shock at x = 0.1 (post-shock water p 1e9, rho 1323.65, u 681.58); z clipped to
               [1e-6, 1-1e-6]; PVRS-HLLC; 2000 steps at CFL 0.5.
"""
from . import _lib as L

Z_EPS = 1e-6


def euler_discs(n=200, n_discs=8, r=(0.03, 0.10), bc=(0, 0, 0, 0), seed=1, limiter=L.VANLEER, ny=None):
    d, p = euler_discs_params(n, n_discs, r, bc, seed, limiter, ny)
    return d, L.random_discs(d, p)


def euler_discs_params(n=200, n_discs=8, r=(0.03, 0.10), bc=(0, 0, 0, 0), seed=1, limiter=L.VANLEER, ny=None):
    """Descriptor and generator parameters of euler_discs (for device-side generation)."""
    ny = n if ny is None else ny
    d = L.make_desc(L.EULER, n, ny, lx=1.0, ly=ny / n, gamma1=1.4, bc=bc, limiter=limiter)
    p = L.DiscParams()
    p.seed, p.n = seed, n_discs
    p.cx_lo, p.cx_hi, p.cy_lo, p.cy_hi = 0.1, 0.9, 0.1 * (ny / n), 0.9 * (ny / n)
    p.r_lo, p.r_hi = r
    p.p_lo, p.p_hi, p.rho_lo, p.rho_hi = 5.0, 20.0, 0.5, 2.0
    for i, v in enumerate((1.0, 0.0, 0.0, 1.0)):
        p.ambient[i] = v
    return d, p


def euler_200(seed=1):
    return euler_discs(200, 8, (0.03, 0.10), (0, 0, 0, 0), seed)


def euler_2048(seed=1):
    return euler_discs(2048, 64, (0.005, 0.03), (1, 1, 0, 0), seed)


def euler_2048_godunov(seed=1):
    """euler_2048 with the first-order Godunov flux (ghost width 1, as the reference sets it)."""
    d, p = euler_discs_params(2048, 64, (0.005, 0.03), (1, 1, 0, 0), seed)
    dg = L.make_desc(L.EULER, d.nx, d.ny, lx=1.0, ly=1.0, gamma1=1.4, bc=(1, 1, 0, 0), limiter=d.limiter,
                     flux_solver=L.GODUNOV)
    return dg, L.random_discs(dg, p)


def euler_slab(rows=256, index=3, seed=1):
    """One GPU's share of euler_2048 split into 2048/rows y-slabs of full rows: rows
    [index*rows, (index+1)*rows) of its initial state, as a 2048 x rows grid with the same cell
    size. Its bottom and top sides are transmissive here; in a sharded run they would read the
    neighbouring slabs' halo rows instead, which costs the sweeps the same work."""
    d, h = euler_2048(seed)
    ng = d.ng
    ds = L.make_desc(L.EULER, d.nx, rows, lx=1.0, ly=rows / d.nx, gamma1=1.4, bc=(1, 1, 0, 0), limiter=d.limiter)
    y0 = index * rows
    return ds, h[y0:y0 + rows + 2 * ng].copy()


def allaire_bubbles(n=1024, n_bubbles=16, seed=1, ny=None, wave_speeds=L.PVRS, limiter=L.MINMOD):
    d, p = allaire_bubbles_params(n, n_bubbles, seed, ny, wave_speeds, limiter)
    return d, L.shocked_bubbles(d, p)


def allaire_bubbles_params(n=1024, n_bubbles=16, seed=1, ny=None, wave_speeds=L.PVRS, limiter=L.MINMOD):
    """Descriptor and generator parameters of allaire_bubbles (for device-side generation)."""
    ny = n if ny is None else ny
    d = L.make_desc(L.ALLAIRE, n, ny, lx=1.0, ly=ny / n, gamma1=4.4, pinf1=6e8, gamma2=1.4, pinf2=0.0,
                    bc=(0, 0, 0, 0), limiter=limiter, wave_speeds=wave_speeds, z_lo=Z_EPS, z_hi=1.0 - Z_EPS)
    zw, za = 1.0 - Z_EPS, Z_EPS
    p = L.BubbleParams()
    p.seed, p.n = seed, n_bubbles
    p.cx_lo, p.cx_hi, p.cy_lo, p.cy_hi = 0.15, 0.95, 0.05 * (ny / n), 0.95 * (ny / n)
    p.r_lo, p.r_hi = 0.02, 0.06
    water = (zw * 1000.0, (1 - zw) * 1.0, 0.0, 0.0, 1e5, zw)
    air = (za * 1000.0, (1 - za) * 1.0, 0.0, 0.0, 1e5, za)
    post = (zw * 1323.65, (1 - zw) * 1.0, 681.58, 0.0, 1e9, zw)
    for i in range(6):
        p.water[i], p.air[i], p.postshock[i] = water[i], air[i], post[i]
    p.shock_x = 0.1
    return d, p


def allaire_1024(seed=1):
    return allaire_bubbles(1024, 16, seed)


def batched_mix(n_grids=256, seed=2024, lo=128, hi=1024, euler_frac=0.75):
    """The batched workload: n_grids independent grids, exactly round(euler_frac * n_grids) of them
    Euler and the rest Allaire (the models in a seeded random order; 192 and 64 of 256), with nx
    and ny drawn independently log-uniform in [lo, hi] (square cells, domain [0,1] x [0,ny/nx]);
    Euler grids carry 8 seeded discs (radii U[0.03,0.10]), Allaire grids 16 seeded air bubbles
    behind a planar shock. Returns [(desc, state, cfl)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n_euler = int(round(euler_frac * n_grids))
    is_euler = rng.permutation(np.arange(n_grids) < n_euler)
    out = []
    for k in range(n_grids):
        nx = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        ny = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        if is_euler[k]:
            d, h = euler_discs(nx, 8, (0.03, 0.10), (0, 0, 0, 0), seed=seed * 1000 + k, ny=ny)
            out.append((d, h, 0.8))
        else:
            d, h = allaire_bubbles(nx, 16, seed=seed * 1000 + k, ny=ny)
            out.append((d, h, 0.5))
    return out
