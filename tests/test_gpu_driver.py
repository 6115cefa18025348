"""The settings-file driver (cfd-art_amd/cfdart_run) end to end on the GPU: it reads the
reference's settings format, runs the device loop, and writes the reference's gnuplot frames
(state, schlieren) and mass-change file. Checked: file shapes and invariants, the Allaire
form of Toro's test 1 against the exact Riemann solution, and the fixed-step mass history."""
import glob
import os
import subprocess

import numpy as np
import pytest

from exact_riemann import Side, solve

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cfd-art_amd", "cfdart_run")


def run_driver(tmp_path, **kv):
    settings = tmp_path / "settings.txt"
    kv.setdefault("outputpath", str(tmp_path) + "/")
    settings.write_text("".join(f"{k} {v}\n" for k, v in kv.items()))
    print(f"case driver {kv}", flush=True)
    r = subprocess.run([EXE, str(settings)], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def read_frame(path, ncols):
    rows = [l.split() for l in open(path) if l.strip()]
    a = np.array(rows, dtype=float)
    assert a.shape[1] == ncols
    return a


@pytest.mark.parametrize("flux", ["MUSCL", "Godunov"])
def test_driver_fixed_steps_frames_and_masses(gpu, tmp_path, flux):
    run_driver(tmp_path, Nx=64, Ny=48, CFL=0.8, splitting_scheme="alternating_splitting", test_case="random_discs",
               output_freq=0, flux_solver=flux, riemann_solver="HLLC", problem="euler", limiter="vanleer",
               num_steps=20, frame_every_steps=10, seed=3, n_discs=5, bc_left="reflective", bc_right="reflective",
               bc_bottom="reflective", bc_top="reflective")
    base = str(tmp_path / f"euler-random_discs-{flux}-64-48")
    states = sorted(glob.glob(base + "-state-*.dat"))
    schl = sorted(glob.glob(base + "-schlieren-*.dat"))
    assert len(states) == 3 and len(schl) == 3  # n = 0, 10, 20
    for p in states:
        a = read_frame(p, 7)  # x y rho u v e p
        assert a.shape[0] == 64 * 48 and np.all(a[:, 2] > 0) and np.all(a[:, 6] > 0)
    for p in schl:
        s = read_frame(p, 3)
        assert s.shape[0] == 64 * 48 and s[:, 2].max() == 1.0 and s[:, 2].min() >= 0.0
    m = np.loadtxt(base + "-masschange.dat")
    assert m.shape == (20, 2) and m[0, 1] == 0.0
    # closed box: the mass change stays at round-off (printed with 6 significant digits)
    assert np.all(np.abs(m[:, 1]) < 1e-9)


@pytest.mark.parametrize("flux,tol", [("MUSCL", 0.012), ("Godunov", 0.03)])
def test_driver_allaire_ttc1_against_exact(gpu, tmp_path, flux, tol):
    run_driver(tmp_path, Nx=200, Ny=2, CFL=0.5, splitting_scheme="naive_splitting", test_case="TTC1_x",
               output_freq=0, flux_solver=flux, riemann_solver="HLLC", problem="allaire")
    base = str(tmp_path / f"allaire-TTC1_x-{flux}-200-2")
    final = [p for p in glob.glob(base + "-state-*.dat") if p.endswith("-state-0.250000.dat")]
    assert final, glob.glob(base + "*")
    a = read_frame(final[0], 8)  # x y rho u v e z p
    row = a[a[:, 1] == a[0, 1]]
    xs = row[:, 0]
    ex = np.array([r[0] for r in solve(Side(1.0, 0.0, 1.0, 1.4, 0.0), Side(0.125, 0.0, 0.1, 1.4, 0.0), 0.5, 0.25,
                                       list(xs))])
    err = np.mean(np.abs(row[:, 2] - ex)) / np.mean(ex)
    print(f"TTC1 {flux}: relative L1 density error {err:.4f}", flush=True)
    assert err < tol
