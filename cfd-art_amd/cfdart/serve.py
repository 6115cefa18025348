"""Bindings of the serving loop (include/cfdart_serve.h): request traces, the continuous-
batching server and the one-at-a-time comparison, plus latency statistics."""
import ctypes as C

import numpy as np

from ._lib import BubbleParams, CfdartError, Desc, DiscParams, lib


class Request(C.Structure):
    _fields_ = [("arrival", C.c_double), ("model", C.c_int), ("n", C.c_int), ("steps", C.c_int),
                ("seed", C.c_ulonglong)]


class ServeOpts(C.Structure):
    _fields_ = [("mem_budget", C.c_size_t), ("capacity", C.c_int), ("chunk_steps", C.c_int),
                ("frame_every", C.c_int), ("frame_host_bytes", C.c_size_t), ("realtime", C.c_int),
                ("y_strip_rows", C.c_int * 2), ("x_rows_per_wave", C.c_int * 2),
                ("plain_y_strip_rows", C.c_int * 2), ("plain_x_rows_per_wave", C.c_int * 2),
                ("live_cells", C.c_longlong), ("arith", C.c_int), ("progress_every", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("admitted", C.c_double), ("finished", C.c_double), ("service", C.c_double),
                ("steps_done", C.c_int), ("frames", C.c_int), ("status", C.c_int), ("checksum", C.c_ulonglong)]


class ServeSummary(C.Structure):
    _fields_ = [("makespan", C.c_double), ("cell_updates", C.c_double), ("frame_bytes", C.c_double),
                ("frames", C.c_longlong), ("chunks", C.c_longlong), ("max_live", C.c_int),
                ("peak_state_bytes", C.c_double)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    P = C.POINTER
    L.cfdart_serve_trace.argtypes = [C.c_ulonglong, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                     C.c_int, P(Request), C.c_int, P(C.c_int)]
    L.cfdart_request_setup.argtypes = [P(Request), P(Desc), P(DiscParams), P(BubbleParams), P(C.c_double)]
    for f in (L.cfdart_serve_batched, L.cfdart_serve_one_at_a_time):
        f.argtypes = [P(Request), C.c_int, P(ServeOpts), P(RequestResult), P(ServeSummary)]
    _bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise CfdartError(rc, f"{what}: {lib().cfdart_last_error().decode()}")


def trace(seed=7, rate=50.0, duration=120.0, euler_frac=0.7, n_lo=64, n_hi=2048, steps_lo=50, steps_hi=2000,
          max_out=100000):
    """The Poisson request trace (cfdart_serve_trace); returns a ctypes array of Request."""
    L = _bind()
    buf = (Request * max_out)()
    n = C.c_int()
    _check(L.cfdart_serve_trace(seed, rate, duration, euler_frac, n_lo, n_hi, steps_lo, steps_hi, buf, max_out,
                                C.byref(n)), "cfdart_serve_trace")
    out = (Request * n.value)()
    C.memmove(out, buf, C.sizeof(Request) * n.value)
    return out


def requests(items):
    """ctypes array of Request from (arrival, model, n, steps, seed) tuples."""
    out = (Request * len(items))()
    for r, (a, m, n, s, seed) in zip(out, items):
        r.arrival, r.model, r.n, r.steps, r.seed = a, m, n, s, seed
    return out


def setup(req):
    """(desc, params, cfl) of a request: DiscParams for Euler, BubbleParams for Allaire."""
    L = _bind()
    d, dp, bp, cfl = Desc(), DiscParams(), BubbleParams(), C.c_double()
    _check(L.cfdart_request_setup(C.byref(req), C.byref(d), C.byref(dp), C.byref(bp), C.byref(cfl)),
           "cfdart_request_setup")
    return d, (dp if req.model == 0 else bp), cfl.value


def opts(mem_budget=200 * 10**9, capacity=4096, chunk_steps=10, frame_every=50, frame_host_bytes=1 << 30,
         realtime=True, tile=None, plain_tile=None, live_cells=0, arith=0, progress_every=0.0):
    o = ServeOpts()
    o.live_cells = int(live_cells)
    o.arith = int(arith)
    o.progress_every = float(progress_every)
    o.mem_budget, o.capacity, o.chunk_steps, o.frame_every = int(mem_budget), capacity, chunk_steps, frame_every
    o.frame_host_bytes, o.realtime = int(frame_host_bytes), 1 if realtime else 0
    for m in (0, 1):
        if tile and m in tile:
            o.y_strip_rows[m], o.x_rows_per_wave[m] = tile[m]
        if plain_tile and m in plain_tile:
            o.plain_y_strip_rows[m], o.plain_x_rows_per_wave[m] = plain_tile[m]
    return o


def _run(fn, name, reqs, o):
    L = _bind()
    res = (RequestResult * len(reqs))()
    s = ServeSummary()
    _check(getattr(L, fn)(reqs, len(reqs), C.byref(o), res, C.byref(s)), name)
    return list(res), s


def serve_batched(reqs, o):
    return _run("cfdart_serve_batched", "cfdart_serve_batched", reqs, o)


def serve_one_at_a_time(reqs, o):
    return _run("cfdart_serve_one_at_a_time", "cfdart_serve_one_at_a_time", reqs, o)


def latency_stats(reqs, results):
    """p50 / p99 / mean of finished - arrival over the requests that completed (seconds)."""
    lat = np.array([r.finished - q.arrival for q, r in zip(reqs, results) if r.status == 0])
    if lat.size == 0:
        return {"p50": None, "p99": None, "mean": None, "n": 0}
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "mean": float(lat.mean()),
            "n": int(lat.size)}
