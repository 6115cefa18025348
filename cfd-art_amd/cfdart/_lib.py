"""ctypes declarations of include/cfdart.h and include/cfdart_host.h."""
import ctypes as C
import math
import os

import numpy as np

EULER, ALLAIRE = 0, 1
TRANSMISSIVE, REFLECTIVE = 0, 1
MINMOD, VANLEER = 0, 1
GARRICK, PVRS = 0, 1
NAIVE, STRANG, ALTERNATING = 0, 1, 2
MUSCL_HANCOCK, GODUNOV = 0, 1
Y_TRANSPOSED = -1  # set_tile: the transposed y tile, 16 columns (CFDART_Y_TRANSPOSED)
Y_TRANSPOSED_NARROW = -2  # 8 columns (CFDART_Y_TRANSPOSED_NARROW)
ARITH_EXACT, ARITH_FAST = 0, 1  # set_arith (CFDART_ARITH_EXACT, CFDART_ARITH_FAST)

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(os.path.dirname(_HERE), "libcfdart.so")


class CfdartError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cfdart error {code}: {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("gamma1", C.c_double), ("pinf1", C.c_double), ("gamma2", C.c_double), ("pinf2", C.c_double),
        ("bc", C.c_int * 4), ("limiter", C.c_int), ("wave_speeds", C.c_int),
        ("colsweep_uses_dx", C.c_int), ("z_lo", C.c_double), ("z_hi", C.c_double),
        ("flux_solver", C.c_int),
    ]


class DiscParams(C.Structure):
    _fields_ = [
        ("seed", C.c_ulonglong), ("n", C.c_int),
        ("cx_lo", C.c_double), ("cx_hi", C.c_double), ("cy_lo", C.c_double), ("cy_hi", C.c_double),
        ("r_lo", C.c_double), ("r_hi", C.c_double), ("p_lo", C.c_double), ("p_hi", C.c_double),
        ("rho_lo", C.c_double), ("rho_hi", C.c_double), ("ambient", C.c_double * 4),
    ]


class BubbleParams(C.Structure):
    _fields_ = [
        ("seed", C.c_ulonglong), ("n", C.c_int),
        ("cx_lo", C.c_double), ("cx_hi", C.c_double), ("cy_lo", C.c_double), ("cy_hi", C.c_double),
        ("r_lo", C.c_double), ("r_hi", C.c_double),
        ("water", C.c_double * 6), ("air", C.c_double * 6), ("postshock", C.c_double * 6),
        ("shock_x", C.c_double),
    ]


_lib = None


def lib():
    """Load libcfdart.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise CfdartError(-5, f"{library_path} not built (run __graft_entry__.build())")
        L = C.CDLL(library_path)
        P = C.POINTER
        dp = P(C.c_double)
        L.cfdart_last_error.restype = C.c_char_p
        L.cfdart_device_count.restype = C.c_int
        L.cfdart_create.argtypes = [P(Desc), P(C.c_void_p)]
        L.cfdart_destroy.argtypes = [C.c_void_p]
        L.cfdart_destroy.restype = None
        L.cfdart_upload.argtypes = [C.c_void_p, dp]
        L.cfdart_download.argtypes = [C.c_void_p, dp]
        L.cfdart_sweep.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.cfdart_compute_dt.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
        L.cfdart_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.cfdart_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, dp, P(C.c_int)]
        L.cfdart_run_mass.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, dp, P(C.c_int), dp]
        L.cfdart_set_time.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.cfdart_stats.argtypes = [C.c_void_p, P(C.c_longlong)]
        L.cfdart_sync.argtypes = [C.c_void_p]
        L.cfdart_stream.argtypes = [C.c_void_p]
        L.cfdart_stream.restype = C.c_void_p
        L.cfdart_set_strip_rows.argtypes = [C.c_void_p, C.c_int]
        L.cfdart_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.cfdart_profile.argtypes = [C.c_void_p, dp, P(C.c_longlong)]
        L.cfdart_ic_random_discs.argtypes = [P(Desc), P(DiscParams), dp]
        L.cfdart_ic_shocked_bubbles.argtypes = [P(Desc), P(BubbleParams), dp]
        L.cfdart_ic_planar.argtypes = [P(Desc), dp, dp, C.c_double, C.c_double, C.c_double, dp]
        L.cfdart_prim_to_cons.argtypes = [P(Desc), dp, dp]
        L.cfdart_copy_state.argtypes = [C.c_void_p, C.c_int]
        L.cfdart_run_noalloc.argtypes = [P(Desc), C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p, P(C.c_double),
                                         P(C.c_int)]
        L.cfdart_test_case.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_char_p, P(Desc), dp,
                                       C.c_size_t, P(C.c_double)]
        L.cfdart_set_tile.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.cfdart_set_arith.argtypes = [C.c_void_p, C.c_int]
        L.cfdart_init_random_discs.argtypes = [C.c_void_p, P(DiscParams)]
        L.cfdart_init_shocked_bubbles.argtypes = [C.c_void_p, P(BubbleParams)]
        L.cfdart_checksum.argtypes = [C.c_void_p, P(C.c_ulonglong)]
        L.cfdart_frame.argtypes = [C.c_void_p, C.c_void_p]
        L.cfdart_schlieren.argtypes = [C.c_void_p, dp, dp]
        # batch (include/cfdart_batch.h)
        L.cfdart_batch_create.argtypes = [C.c_int, C.c_size_t, P(C.c_void_p)]
        L.cfdart_batch_destroy.argtypes = [C.c_void_p]
        L.cfdart_batch_destroy.restype = None
        L.cfdart_batch_add.argtypes = [C.c_void_p, P(Desc), dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_int, C.c_int, P(C.c_int)]
        L.cfdart_batch_add_generated.argtypes = [C.c_void_p, P(Desc), P(DiscParams), P(BubbleParams), C.c_int,
                                                 C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, P(C.c_int)]
        L.cfdart_batch_poll.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
        L.cfdart_batch_checksum.argtypes = [C.c_void_p, C.c_int, P(C.c_ulonglong)]
        L.cfdart_batch_checksum_async.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.cfdart_batch_remove.argtypes = [C.c_void_p, C.c_int]
        L.cfdart_batch_step.argtypes = [C.c_void_p, C.c_int]
        L.cfdart_batch_download.argtypes = [C.c_void_p, C.c_int, dp]
        L.cfdart_batch_status.argtypes = [C.c_void_p, C.c_int, dp, P(C.c_int), P(C.c_int)]
        L.cfdart_batch_stats.argtypes = [C.c_void_p, C.c_int, P(C.c_longlong)]
        L.cfdart_batch_sync.argtypes = [C.c_void_p]
        L.cfdart_batch_set_tile.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.cfdart_batch_set_arith.argtypes = [C.c_void_p, C.c_int]
        L.cfdart_batch_info.argtypes = [C.c_void_p, P(C.c_int), P(C.c_size_t), P(C.c_int)]
        L.cfdart_batch_frame.argtypes = [C.c_void_p, C.c_int, dp]
        L.cfdart_batch_frames.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_int)]
        L.cfdart_batch_frames_done.argtypes = [C.c_void_p]
        L.cfdart_batch_stream.argtypes = [C.c_void_p]
        L.cfdart_batch_stream.restype = C.c_void_p
        L.cfdart_host_alloc.argtypes = [C.c_size_t]
        L.cfdart_host_alloc.restype = C.c_void_p
        L.cfdart_host_free.argtypes = [C.c_void_p]
        L.cfdart_host_free.restype = None
        L.cfdart_time_tile.argtypes = [P(Desc), dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CfdartError(rc, lib().cfdart_last_error().decode())
    return rc


def _dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def device_count():
    return lib().cfdart_device_count()


def ncomp(model):
    return 6 if model == ALLAIRE else 4


def make_desc(model=EULER, nx=200, ny=200, x0=0.0, y0=0.0, lx=1.0, ly=1.0, gamma1=1.4, pinf1=0.0,
              gamma2=1.4, pinf2=0.0, bc=(0, 0, 0, 0), limiter=MINMOD, wave_speeds=GARRICK,
              colsweep_uses_dx=1, z_lo=0.0, z_hi=1.0, flux_solver=0, ng=None):
    """flux_solver: MUSCL_HANCOCK (ghost width 2) or GODUNOV (ghost width 1, as the reference)."""
    d = Desc()
    d.model, d.nx, d.ny = model, nx, ny
    d.ng = ng if ng is not None else (1 if flux_solver == GODUNOV else 2)
    d.flux_solver = flux_solver
    d.x0, d.y0, d.dx, d.dy = x0, y0, lx / nx, ly / ny
    d.gamma1, d.pinf1, d.gamma2, d.pinf2 = gamma1, pinf1, gamma2, pinf2
    for i in range(4):
        d.bc[i] = bc[i]
    d.limiter, d.wave_speeds, d.colsweep_uses_dx = limiter, wave_speeds, colsweep_uses_dx
    d.z_lo, d.z_hi = z_lo, z_hi
    return d


def empty_state(d):
    return np.zeros((d.ny + 2 * d.ng, d.nx + 2 * d.ng, ncomp(d.model)), dtype=np.float64)


def test_case(problem, name, nx, ny, flux_solver="MUSCL"):
    """A named test case of the settings-file driver (cfdart_test_case): (desc, state, T)."""
    d = Desc()
    T = C.c_double()
    _check(lib().cfdart_test_case(problem.encode(), name.encode(), nx, ny, flux_solver.encode(), C.byref(d), None,
                                  0, C.byref(T)))
    h = empty_state(d)
    _check(lib().cfdart_test_case(problem.encode(), name.encode(), nx, ny, flux_solver.encode(), C.byref(d),
                                  _dptr(h), h.size, C.byref(T)))
    return d, h, T.value


def random_discs(d, p):
    h = empty_state(d)
    _check(lib().cfdart_ic_random_discs(C.byref(d), C.byref(p), _dptr(h)))
    return h


def shocked_bubbles(d, p):
    h = empty_state(d)
    _check(lib().cfdart_ic_shocked_bubbles(C.byref(d), C.byref(p), _dptr(h)))
    return h


def planar(d, w_under, w_over, a, b, c):
    h = empty_state(d)
    wu = np.ascontiguousarray(w_under, dtype=np.float64)
    wo = np.ascontiguousarray(w_over, dtype=np.float64)
    _check(lib().cfdart_ic_planar(C.byref(d), _dptr(wu), _dptr(wo), a, b, c, _dptr(h)))
    return h


def prim_to_cons(d, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    u = np.zeros(ncomp(d.model))
    _check(lib().cfdart_prim_to_cons(C.byref(d), _dptr(w), _dptr(u)))
    return u


class Grid:
    """An opaque device-resident grid (cfdart_grid*)."""

    def __init__(self, desc):
        self.desc = desc
        h = C.c_void_p()
        _check(lib().cfdart_create(C.byref(desc), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().cfdart_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        _check(lib().cfdart_upload(self.h, _dptr(state)))

    def download(self, out=None):
        if out is None:
            out = empty_state(self.desc)
        _check(lib().cfdart_download(self.h, _dptr(out)))
        return out

    def sweep(self, axis, dt):
        _check(lib().cfdart_sweep(self.h, axis, dt))

    def step(self, splitting, n, dt):
        _check(lib().cfdart_step(self.h, splitting, n, dt))

    def compute_dt(self, n, t, T, cfl):
        dt = C.c_double()
        mass = (C.c_double * 2)()
        _check(lib().cfdart_compute_dt(self.h, n, t, T, cfl, C.byref(dt), mass))
        return dt.value, (mass[0], mass[1])

    def set_time(self, t, n):
        _check(lib().cfdart_set_time(self.h, t, n))

    def run(self, splitting, max_steps, T=math.inf, cfl=0.8, sync=True):
        if not sync:
            _check(lib().cfdart_run(self.h, splitting, max_steps, T, cfl, None, None))
            return None
        t = C.c_double()
        n = C.c_int()
        _check(lib().cfdart_run(self.h, splitting, max_steps, T, cfl, C.byref(t), C.byref(n)))
        return t.value, n.value

    def init_generated(self, params):
        """Device-side initial state from DiscParams (Euler) or BubbleParams (Allaire)."""
        if isinstance(params, DiscParams):
            _check(lib().cfdart_init_random_discs(self.h, C.byref(params)))
        else:
            _check(lib().cfdart_init_shocked_bubbles(self.h, C.byref(params)))

    def checksum(self):
        v = C.c_ulonglong()
        _check(lib().cfdart_checksum(self.h, C.byref(v)))
        return v.value

    def schlieren(self):
        """(image [ny, nx], max |grad rho|) of the current state, computed on the device."""
        img = np.empty((self.desc.ny, self.desc.nx), dtype=np.float64)
        m = C.c_double()
        _check(lib().cfdart_schlieren(self.h, _dptr(img), C.byref(m)))
        return img, m.value

    def run_mass(self, splitting, max_steps, T=math.inf, cfl=0.8):
        """run() that also returns the mass history: array [steps taken, 3] of (t, mass1, mass2)."""
        t = C.c_double()
        n = C.c_int()
        hist = np.zeros((max(max_steps, 1), 3))
        n0 = self.run(splitting, 0, T, cfl)[1] if max_steps else 0
        _check(lib().cfdart_run_mass(self.h, splitting, max_steps, T, cfl, C.byref(t), C.byref(n), _dptr(hist)))
        return t.value, n.value, hist[: n.value - n0]

    def stats(self):
        c = (C.c_longlong * 4)()
        _check(lib().cfdart_stats(self.h, c))
        return list(c)

    def sync(self):
        _check(lib().cfdart_sync(self.h))

    def stream(self):
        return lib().cfdart_stream(self.h)

    def set_profiling(self, enable):
        _check(lib().cfdart_set_profiling(self.h, 1 if enable else 0))

    def profile(self):
        ms = (C.c_double * 2)()
        n = (C.c_longlong * 2)()
        _check(lib().cfdart_profile(self.h, ms, n))
        return {"x": (ms[0], n[0]), "y": (ms[1], n[1])}

    def set_tile(self, y_strip_rows, x_rows_per_wave):
        _check(lib().cfdart_set_tile(self.h, y_strip_rows, x_rows_per_wave))

    def set_arith(self, mode):
        """ARITH_EXACT (default: bitwise the reference's arithmetic) or ARITH_FAST."""
        _check(lib().cfdart_set_arith(self.h, mode))

    def set_strip_rows(self, rows):
        _check(lib().cfdart_set_strip_rows(self.h, rows))


class Batch:
    """Many grids advancing together (include/cfdart_batch.h)."""

    def __init__(self, capacity, mem_budget_bytes=0):
        h = C.c_void_p()
        _check(lib().cfdart_batch_create(capacity, mem_budget_bytes, C.byref(h)))
        self.h = h
        self.descs = {}

    def close(self):
        if self.h:
            lib().cfdart_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, desc, state, splitting=ALTERNATING, max_steps=1, T=math.inf, cfl=0.8, y_strip_rows=0,
            x_rows_per_wave=0):
        """Returns the slot id; raises CfdartError (code -2) when admission is refused."""
        state = np.ascontiguousarray(state, dtype=np.float64)
        slot = C.c_int()
        _check(lib().cfdart_batch_add(self.h, C.byref(desc), _dptr(state), splitting, max_steps, T, cfl,
                                      y_strip_rows, x_rows_per_wave, C.byref(slot)))
        self.descs[slot.value] = desc
        return slot.value

    def add_generated(self, desc, params, splitting=ALTERNATING, max_steps=1, T=math.inf, cfl=0.8,
                      y_strip_rows=0, x_rows_per_wave=0):
        """Admits a grid whose initial state the device generates from DiscParams (Euler) or
        BubbleParams (Allaire). Returns the slot id; CfdartError (code -2) when refused."""
        slot = C.c_int()
        dp = C.byref(params) if isinstance(params, DiscParams) else None
        bp = C.byref(params) if isinstance(params, BubbleParams) else None
        _check(lib().cfdart_batch_add_generated(self.h, C.byref(desc), dp, bp, splitting, max_steps, T, cfl,
                                                y_strip_rows, x_rows_per_wave, C.byref(slot)))
        self.descs[slot.value] = desc
        return slot.value

    def poll(self, capacity):
        steps = (C.c_int * capacity)()
        flags = (C.c_int * capacity)()
        _check(lib().cfdart_batch_poll(self.h, steps, flags))
        return list(steps), list(flags)

    def checksum(self, slot):
        v = C.c_ulonglong()
        _check(lib().cfdart_batch_checksum(self.h, slot, C.byref(v)))
        return v.value

    def try_add(self, *a, **k):
        try:
            return self.add(*a, **k)
        except CfdartError as e:
            if e.code == -2:
                return None
            raise

    def remove(self, slot):
        _check(lib().cfdart_batch_remove(self.h, slot))
        self.descs.pop(slot, None)

    def step(self, nsteps=1):
        _check(lib().cfdart_batch_step(self.h, nsteps))

    def sync(self):
        _check(lib().cfdart_batch_sync(self.h))

    def download(self, slot):
        out = empty_state(self.descs[slot])
        _check(lib().cfdart_batch_download(self.h, slot, _dptr(out)))
        return out

    def status(self, slot):
        t = C.c_double()
        n = C.c_int()
        a = C.c_int()
        _check(lib().cfdart_batch_status(self.h, slot, C.byref(t), C.byref(n), C.byref(a)))
        return t.value, n.value, bool(a.value)

    def stats(self, slot):
        c = (C.c_longlong * 4)()
        _check(lib().cfdart_batch_stats(self.h, slot, c))
        return list(c)

    def set_tile(self, slot, y_strip_rows, x_rows_per_wave):
        _check(lib().cfdart_batch_set_tile(self.h, slot, y_strip_rows, x_rows_per_wave))

    def set_arith(self, mode):
        """ARITH_EXACT (default) or ARITH_FAST for every grid of the batch, from the next step."""
        _check(lib().cfdart_batch_set_arith(self.h, mode))

    def info(self):
        live = C.c_int()
        b = C.c_size_t()
        t = C.c_int()
        _check(lib().cfdart_batch_info(self.h, C.byref(live), C.byref(b), C.byref(t)))
        return live.value, b.value, t.value

    def frame(self, slot, host_ptr):
        _check(lib().cfdart_batch_frame(self.h, slot, C.cast(host_ptr, C.POINTER(C.c_double))))

    def frames(self, slots, host_ptrs):
        """Frames of several grids in one group; returns each frame's status (0 taken,
        -2 skipped: the staging ring could not grow within the memory budget)."""
        n = len(slots)
        s = (C.c_int * n)(*slots)
        p = (C.c_void_p * n)(*host_ptrs)
        st = (C.c_int * n)()
        _check(lib().cfdart_batch_frames(self.h, n, s, p, st))
        return list(st)

    def frames_done(self):
        _check(lib().cfdart_batch_frames_done(self.h))


def time_tile(desc, y_strip_rows, x_rows_per_wave, reps=5, state=None, arith=ARITH_EXACT):
    ux = C.c_double()
    uy = C.c_double()
    sp = None if state is None else _dptr(np.ascontiguousarray(state, dtype=np.float64))
    _check(lib().cfdart_time_tile(C.byref(desc), sp, y_strip_rows, x_rows_per_wave, reps, arith, C.byref(ux),
                                  C.byref(uy)))
    return ux.value, uy.value


def state_checksum(state, ng=2):
    """Host restatement of the device state checksum (cfdart_checksum) for a host state in the
    cfdart.h layout: sum mod 2^64 of mix64(bits ^ (index * 0x9e3779b97f4a7c15)) over the
    interior values in [component][y][x] order, mix64 being the splitmix64 finaliser."""
    a = np.ascontiguousarray(np.moveaxis(state[ng:-ng, ng:-ng, :], 2, 0)).view(np.uint64).ravel()
    idx = np.arange(a.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = a ^ (idx * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return int(np.sum(z, dtype=np.uint64))
