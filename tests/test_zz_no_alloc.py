"""cfdart_run_noalloc, the time loop that allocates nothing (its control block and reduction
partials live in the library's static device array), against cfdart_run on a handle: two calls
in a row, the second on a smaller grid, each bitwise equal in state, t and n. The caller's
device buffers are allocated by the test through the HIP runtime the library itself links."""
import ctypes as C
import math

import numpy as np
import pytest

import cfdart
from util import bitwise_equal, interior

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libcfdart.so is linked against."""

    def __init__(self):
        import subprocess
        out = subprocess.run(["ldd", cfdart.library_path], capture_output=True, text=True).stdout
        path = next(l.split("=>")[1].split()[0] for l in out.splitlines() if "libamdhip64" in l)
        self.L = C.CDLL(path)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def h2d(self, dst, a):
        assert self.L.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def free(self, p):
        assert self.L.hipFree(p) == 0


def layout(d, h, pitch):
    """The state in the library's device layout: ncomp planes of pitch*ny doubles, interior only."""
    nc, ng = cfdart.ncomp(d.model), d.ng
    a = np.zeros((nc, d.ny, pitch))
    a[:, :, :d.nx] = np.moveaxis(h[ng:-ng, ng:-ng, :], 2, 0)
    return np.ascontiguousarray(a)


def run_noalloc(hip, d, h, splitting, steps, cfl, arith, pitch):
    a = layout(d, h, pitch)
    state, work = hip.alloc(a.nbytes), hip.alloc(a.nbytes)
    try:
        hip.h2d(state, a)
        t, n = C.c_double(), C.c_int()
        rc = cfdart.lib().cfdart_run_noalloc(C.byref(d), state, work, pitch, arith, splitting, steps, math.inf, cfl,
                                             0.0, 0, None, C.byref(t), C.byref(n))
        assert rc == 0, cfdart.lib().cfdart_last_error()
        out = np.empty_like(a)
        hip.d2h(out, state)
    finally:
        hip.free(state)
        hip.free(work)
    return np.moveaxis(out[:, :, :d.nx], 0, 2), t.value, n.value


def run_handle(d, h, splitting, steps, cfl, arith):
    g = cfdart.Grid(d)
    g.set_arith(arith)
    g.upload(h)
    t, n = g.run(splitting, steps, cfl=cfl)
    out = interior(g.download())
    g.close()
    return out, t, n


def test_run_noalloc_equals_run(gpu):
    hip = Hip()
    calls = [
        (cfdart.configs.euler_discs(160, 6, (0.05, 0.15), (1, 0, 0, 1), seed=4), cfdart.ALTERNATING, 12, 0.8,
         cfdart.ARITH_EXACT, 160),
        (cfdart.configs.allaire_bubbles(64, 3, seed=5, ny=48), cfdart.STRANG, 9, 0.5, cfdart.ARITH_FAST, 67),
    ]
    for (d, h), splitting, steps, cfl, arith, pitch in calls:
        print(f"case noalloc model={d.model} {d.nx}x{d.ny} splitting={splitting} steps={steps} arith={arith} "
              f"pitch={pitch}", flush=True)
        got, t, n = run_noalloc(hip, d, h, splitting, steps, cfl, arith, pitch)
        want, tw, nw = run_handle(d, h, splitting, steps, cfl, arith)
        assert n == nw == steps and t == tw
        assert bitwise_equal(got, want)
