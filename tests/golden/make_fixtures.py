"""Writes the small golden fixtures of tests/golden/: an initial state from the product's seeded
generators and the oracle's state after a few steps. They pin the oracle (and, through the GPU
tests, the kernels) against future drift. Regenerate only on a deliberate numerics change:
    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "cfd-art_amd"), os.path.join(ROOT, "oracle")]

import cfdart  # noqa: E402
import oracle  # noqa: E402

CASES = {
    "euler_48x32_minmod_naive": dict(kind="euler", nx=48, ny=32, limiter=cfdart.MINMOD, bc=(1, 0, 0, 1),
                                     splitting=cfdart.NAIVE, steps=10, cfl=0.8),
    "euler_48x32_vanleer_strang": dict(kind="euler", nx=48, ny=32, limiter=cfdart.VANLEER, bc=(0, 0, 1, 1),
                                       splitting=cfdart.STRANG, steps=10, cfl=0.8),
    "allaire_48x32_pvrs_alt": dict(kind="allaire", nx=48, ny=32, ws=cfdart.PVRS, splitting=cfdart.ALTERNATING,
                                   steps=10, cfl=0.5),
    "allaire_48x32_garrick_naive": dict(kind="allaire", nx=48, ny=32, ws=cfdart.GARRICK, splitting=cfdart.NAIVE,
                                        steps=10, cfl=0.5),
    "euler_48x32_godunov_naive": dict(kind="euler", nx=48, ny=32, limiter=cfdart.MINMOD, bc=(1, 0, 0, 1),
                                      splitting=cfdart.NAIVE, steps=10, cfl=0.8, godunov=True),
    "allaire_48x32_godunov_alt": dict(kind="allaire", nx=48, ny=32, ws=cfdart.PVRS, splitting=cfdart.ALTERNATING,
                                      steps=10, cfl=0.5, godunov=True),
}


def build_case(c):
    if c["kind"] == "euler":
        d, p = cfdart.configs.euler_discs_params(c["nx"], 5, (0.05, 0.2), c["bc"], seed=21, limiter=c["limiter"],
                                                 ny=c["ny"])
    else:
        d, p = cfdart.configs.allaire_bubbles_params(c["nx"], 3, seed=21, ny=c["ny"], wave_speeds=c["ws"])
    if c.get("godunov"):
        d.flux_solver, d.ng = cfdart.GODUNOV, 1
    h = cfdart.random_discs(d, p) if c["kind"] == "euler" else cfdart.shocked_bubbles(d, p)
    return d, h


def desc_array(d):
    return np.array([d.model, d.nx, d.ny, d.ng, d.x0, d.y0, d.dx, d.dy, d.gamma1, d.pinf1, d.gamma2, d.pinf2,
                     d.bc[0], d.bc[1], d.bc[2], d.bc[3], d.limiter, d.wave_speeds, d.colsweep_uses_dx,
                     d.z_lo, d.z_hi, d.flux_solver], dtype=np.float64)


def main():
    force = "--force" in sys.argv[1:]
    for name, c in CASES.items():
        if not force and os.path.exists(os.path.join(HERE, name + ".npz")):
            continue
        d, h = build_case(c)
        s = oracle.Sim(d, h)
        k, hist = s.run(c["splitting"], c["steps"], cfl=c["cfl"])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), desc=desc_array(d), initial=h, final=s.grid,
                            dt=hist, splitting=c["splitting"], steps=k, cfl=c["cfl"],
                            counters=np.array(s.counters()))
        print(name, k, s.t)


if __name__ == "__main__":
    main()
