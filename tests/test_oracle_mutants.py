"""The oracle's branch pins catch plausible mistakes: each case plants one mistake in a copy of
oracle/cfd_oracle.c (a textual substitution), builds it with the oracle's own flags into a
temporary library, and runs the pins named beside it against that build in a separate process
(CFD_ORACLE_LIB). The pins must fail on every mutant and pass on the unmodified source built the
same way. Nothing here touches the oracle's sources or its built library."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "cfd_oracle.c")
CFLAGS = ["-O2", "-fPIC", "-std=c11", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared"]

B = "tests/test_oracle_branches.py"
P = "tests/test_oracle_pins.py"

# (name, [(text, replacement), ...], pins that must fail)
MUTANTS = [
    ("repair rho threshold strict", [("if (rho <= 1e-12) rho = rho_old;", "if (rho < 1e-12) rho = rho_old;")],
     [f"{B}::test_repair_thresholds_inclusive"]),
    ("repair e threshold strict", [("if (e <= 1e-12) e = e_old;", "if (e < 1e-12) e = e_old;")],
     [f"{B}::test_repair_thresholds_inclusive"]),
    ("repair e_old from the new state",
     [("double e_old = old[3] / rho_old - 0.5 * (u_old * u_old + v_old * v_old);",
       "double e_old = fut[3] / rho - 0.5 * (u * u + v * v);")],
     [f"{B}::test_repair_sources"]),
    ("repair velocities from the old state", [("double W[4] = {rho, u, v, p};", "double W[4] = {rho, u_old, v_old, p};")],
     [f"{B}::test_repair_sources"]),
    ("M3 zeroes only the left slope",
     [("for (int k = 0; k < n; ++k) { delL[k] = 0.0; delR[k] = 0.0; }", "for (int k = 0; k < n; ++k) { delL[k] = 0.0; }")],
     [f"{B}::test_slope_fallback_allaire"]),
    ("M5 falls back to s0, s3", [("{ UL[k] = s1[k]; UR[k] = s2[k]; }", "{ UL[k] = s0[k]; UR[k] = s3[k]; }")],
     [f"{B}::test_evolve_fallback_euler"]),
    ("M5 snaps the raw fallback states",
     [("{ UL[k] = s1[k]; UR[k] = s2[k]; }",
       "{ UL[k] = s1[k]; UR[k] = s2[k]; } if (all) { orc_is_physical_snap(d, UL); orc_is_physical_snap(d, UR); }")],
     [f"{B}::test_evolve_fallback_allaire_raw_unsnapped"]),
    ("M5 tests without snapping",
     [("\n    ok = all ? (orc_is_physical_snap(d, UL) && orc_is_physical_snap(d, UR))",
       "\n    ok = all ? (orc_is_physical(d, UL) && orc_is_physical(d, UR))")],
     [f"{B}::test_evolve_snapping_allaire"]),
    ("Strang X(dt) Y(dt) X(dt)",
     [("axes[0] = 0; dts[0] = 0.5 * dt;", "axes[0] = 0; dts[0] = dt;"),
      ("axes[2] = 0; dts[2] = 0.5 * dt;", "axes[2] = 0; dts[2] = dt;")],
     [f"{B}::test_splitting_y_identity_state", f"{B}::test_splitting_composed_from_sweeps"]),
    ("alternating odd steps in the even order",
     [("        axes[0] = 1; dts[0] = dt;\n        axes[1] = 0; dts[1] = dt;",
       "        axes[0] = 0; dts[0] = dt;\n        axes[1] = 1; dts[1] = dt;")],
     [f"{B}::test_alternating_odd_is_transposed_naive", f"{B}::test_splitting_composed_from_sweeps"]),
    ("shock_speed_jump without the 0.5 of the mean pressure",
     [("double p_avg = 0.5 * (pstar + p);", "double p_avg = (pstar + p);")],
     [f"{P}::test_shock_speed_jump_matches_rankine_hugoniot"]),
]


def build(tmp_path, subs):
    src = open(SRC).read()
    for old, new in subs:
        assert src.count(old) == 1, f"mutation site not unique in cfd_oracle.c: {old!r}"
        src = src.replace(old, new)
    c = tmp_path / "cfd_oracle.c"
    c.write_text(src)
    so = tmp_path / "liboracle.so"
    subprocess.run(["gcc", *CFLAGS, "-I", os.path.join(ROOT, "oracle"), "-o", str(so), str(c), "-lm"], check=True)
    return str(so)


def run_pins(lib, tests):
    env = dict(os.environ, CFD_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-2000:]


def test_unmodified_build_passes(tmp_path):
    lib = build(tmp_path, [])
    tests = sorted({t for _, _, ts in MUTANTS for t in ts})
    rc, out = run_pins(lib, tests)
    assert rc == 0, out


@pytest.mark.parametrize("name,subs,tests", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_caught(tmp_path, name, subs, tests):
    lib = build(tmp_path, subs)
    rc, out = run_pins(lib, tests)
    assert rc == 1, f"{name}: the pins did not fail (rc {rc})\n{out}"
