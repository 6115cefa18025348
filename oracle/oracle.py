"""ctypes binding of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY: imported by
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the product."""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CFD_ORACLE_LIB names another build of the oracle (the mutation tests load builds with planted
# mistakes through it, in a separate process)
LIB_PATH = os.environ.get("CFD_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class OrcDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("gamma1", C.c_double), ("pinf1", C.c_double), ("gamma2", C.c_double), ("pinf2", C.c_double),
        ("bc", C.c_int * 4), ("limiter", C.c_int), ("wave_speeds", C.c_int),
        ("colsweep_uses_dx", C.c_int), ("z_lo", C.c_double), ("z_hi", C.c_double),
        ("flux_solver", C.c_int),
    ]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        lp = P(C.c_longlong)
        D = P(OrcDesc)
        L.orc_fill_ghosts.argtypes = [D, dp]
        L.orc_sweep.argtypes = [D, C.c_int, dp, dp, C.c_double, lp]
        L.orc_compute_dt.argtypes = [D, dp, C.c_int, C.c_double, C.c_double, C.c_double, dp]
        L.orc_compute_dt.restype = C.c_double
        L.orc_advance.argtypes = [D, C.c_int, dp, dp, C.c_double, C.c_int, lp]
        L.orc_run.argtypes = [D, C.c_int, dp, dp, C.c_int, C.c_double, C.c_double, dp, C.c_int, dp, lp]
        L.orc_run.restype = C.c_int
        L.orc_c2p.argtypes = [D, dp, dp]
        L.orc_p2c.argtypes = [D, dp, dp]
        L.orc_is_physical.argtypes = [D, dp]
        L.orc_is_physical_snap.argtypes = [D, dp]
        L.orc_flux_phys.argtypes = [D, dp, dp]
        L.orc_hllc.argtypes = [D, dp, dp, dp, dp]
        L.orc_face.argtypes = [D, dp, dp, dp, dp, C.c_double, C.c_double, dp, dp, dp, lp]
        L.orc_jacobian.argtypes = [D, dp, dp]
        L.orc_euler_repair.argtypes = [D, dp, dp]
        L.orc_euler_repair.restype = C.c_int
        L.orc_limit.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_limit.restype = C.c_double
        L.orc_zupdate.argtypes = [D, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_zupdate.restype = C.c_double
        L.orc_mixture_eos.argtypes = [D, C.c_double, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_get_threads.argtypes = []
        L.orc_get_threads.restype = C.c_int
        for f in ("orc_pressure", "orc_specific_ie", "orc_soundspeed"):
            getattr(L, f).argtypes = [D, C.c_double, C.c_double, C.c_double]
            getattr(L, f).restype = C.c_double
        L.orc_shock_speed_jump.argtypes = [D, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_shock_speed_jump.restype = C.c_double
        _L = L
    return _L


def desc_from(d):
    """Copy any structure with the descriptor's field names (e.g. cfdart.Desc) into OrcDesc."""
    o = OrcDesc()
    for name, _ in OrcDesc._fields_:
        if name == "bc":
            for i in range(4):
                o.bc[i] = d.bc[i]
        else:
            setattr(o, name, getattr(d, name))
    return o


def make_desc(**kw):
    o = OrcDesc()
    o.ng = 2
    o.colsweep_uses_dx = 1
    o.z_hi = 1.0
    for k, v in kw.items():
        if k == "bc":
            for i in range(4):
                o.bc[i] = v[i]
        else:
            setattr(o, k, v)
    return o


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Sim:
    """The oracle's two-buffer grid and time loop."""

    def __init__(self, d, state):
        self.d = desc_from(d) if not isinstance(d, OrcDesc) else d
        self.grid = np.ascontiguousarray(state, dtype=np.float64).copy()
        self.future = self.grid.copy()
        self.cnt = (C.c_longlong * 4)()
        self.t = 0.0
        self.n = 0

    def compute_dt(self, T=math.inf, cfl=0.8):
        mass = (C.c_double * 2)()
        dt = lib().orc_compute_dt(C.byref(self.d), _dp(self.grid), self.n, self.t, T, cfl, mass)
        return dt, (mass[0], mass[1])

    def sweep(self, axis, dt):
        lib().orc_fill_ghosts(C.byref(self.d), _dp(self.grid))
        lib().orc_sweep(C.byref(self.d), axis, _dp(self.grid), _dp(self.future), dt, self.cnt)
        self.grid, self.future = self.future, self.grid

    def advance(self, splitting, dt, n=None):
        lib().orc_advance(C.byref(self.d), splitting, _dp(self.grid), _dp(self.future), dt,
                          self.n if n is None else n, self.cnt)

    def run(self, splitting, max_steps, T=math.inf, cfl=0.8, dt_hist=None):
        t = C.c_double(self.t)
        hist = np.zeros(max_steps) if dt_hist is None else dt_hist
        k = lib().orc_run(C.byref(self.d), splitting, _dp(self.grid), _dp(self.future), max_steps, T, cfl,
                          C.byref(t), self.n, _dp(hist), self.cnt)
        self.t = t.value
        self.n += k
        return k, hist[:k]

    def counters(self):
        return list(self.cnt)

    def fill_ghosts(self):
        lib().orc_fill_ghosts(C.byref(self.d), _dp(self.grid))
        return self.grid
