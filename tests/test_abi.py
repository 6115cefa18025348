"""The C-ABI library loads on a CPU-only host and exports every symbol include/*.h declares;
descriptor validation and the host-side initial conditions need no GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import cfdart

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert {"cfdart.h", "cfdart_host.h", "cfdart_batch.h", "cfdart_serve.h"} <= set(headers)
    for h in headers:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cfdart_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    lib = cfdart.lib()
    names = declared_functions()
    assert len(names) > 15
    for n in sorted(names):
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", cfdart.library_path], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (cfdart_\w+)", out))
    assert names <= exported, names - exported


def test_library_is_built_for_gfx950():
    blob = open(cfdart.library_path, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob  # the offload bundle's target id


def test_create_rejects_bad_descriptors():
    lib = cfdart.lib()
    h = C.c_void_p()
    good = cfdart.make_desc(cfdart.EULER, 16, 16)
    bad = []
    d = cfdart.make_desc(cfdart.EULER, 1, 16); bad.append(d)
    d = cfdart.make_desc(cfdart.EULER, 16, 16); d.ng = 1; bad.append(d)
    d = cfdart.make_desc(cfdart.EULER, 16, 16, gamma1=1.0); bad.append(d)
    d = cfdart.make_desc(cfdart.EULER, 16, 16, wave_speeds=cfdart.PVRS); bad.append(d)
    d = cfdart.make_desc(cfdart.ALLAIRE, 16, 16, gamma1=4.4, gamma2=1.4, z_lo=0.5, z_hi=0.4); bad.append(d)
    d = cfdart.make_desc(cfdart.EULER, 16, 16, bc=(0, 2, 0, 0)); bad.append(d)
    for d in bad:
        assert lib.cfdart_create(C.byref(d), C.byref(h)) == -1
        assert lib.cfdart_last_error()
    assert good.ng == 2


def test_random_discs_deterministic_and_bounded():
    d, h1 = cfdart.configs.euler_200(seed=3)
    _, h2 = cfdart.configs.euler_200(seed=3)
    _, h3 = cfdart.configs.euler_200(seed=4)
    assert np.array_equal(h1, h2) and not np.array_equal(h1, h3)
    rho = h1[2:-2, 2:-2, 0]
    p = 0.4 * (h1[2:-2, 2:-2, 3] - 0.5 * (h1[2:-2, 2:-2, 1] ** 2 + h1[2:-2, 2:-2, 2] ** 2) / rho)
    assert rho.min() >= 0.5 - 1e-12 and rho.max() <= 2.0 + 1e-12
    assert p.min() >= 1.0 - 1e-12 and p.max() <= 20.0 + 1e-12
    assert np.count_nonzero(p > 1.0 + 1e-9) > 500  # discs are present
    assert np.all(h1[..., 1] == 0.0) and np.all(h1[..., 2] == 0.0)


def test_disc_fraction_is_subsampled_area():
    """One disc of radius R: the total excess of the blended state equals the 10x10 sampled disc
    area (within the sampling error), i.e. sum f * dx * dy ~ pi R^2."""
    d = cfdart.make_desc(cfdart.EULER, 100, 100)
    p = cfdart.DiscParams()
    p.seed, p.n = 1, 1
    p.cx_lo = p.cx_hi = 0.5
    p.cy_lo = p.cy_hi = 0.5
    p.r_lo = p.r_hi = 0.2
    p.p_lo = p.p_hi = 2.0
    p.rho_lo = p.rho_hi = 3.0
    for i, v in enumerate((1.0, 0.0, 0.0, 1.0)):
        p.ambient[i] = v
    h = cfdart.random_discs(d, p)
    f = (h[2:-2, 2:-2, 0] - 1.0) / 2.0
    assert abs(f.sum() * d.dx * d.dy - np.pi * 0.04) < 2e-4


def test_shocked_bubbles_state():
    d, h = cfdart.configs.allaire_bubbles(128, 16, seed=1)
    z = h[2:-2, 2:-2, 5]
    assert z.min() >= 1e-6 - 1e-18 and z.max() <= 1 - 1e-6 + 1e-18
    assert np.count_nonzero(z < 0.5) > 200  # bubbles
    xc = (np.arange(128) + 0.5) / 128
    u = h[2:-2, 2:-2, 2] / (h[2:-2, 2:-2, 0] + h[2:-2, 2:-2, 1])
    assert np.allclose(u[:, xc <= 0.1], 681.58) and np.all(u[:, xc > 0.1] == 0.0)
