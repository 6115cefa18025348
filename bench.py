"""Benchmark: FP64 2D Euler MUSCL-Hancock on one MI355X, 2048x2048.

Workload "euler_2048": 64 seeded high-pressure discs (radii U[0.005,0.03]) on an ambient gas,
reflective left/right, transmissive bottom/top, van Leer + HLLC, alternating x/y splitting,
CFL 0.8 (the first 5 steps at 0.2, as the reference's compute_dt does). A step is one full split
time step (an x and a y sweep + the CFL reduction) on the device-resident grid; the state is
uploaded before the timed region, dt never leaves the device.

Single-grid workloads are timed in both arithmetic modes of the library: the exact one (every
division and square root correctly rounded, bitwise the reference's arithmetic) and the fast one
(divisions and square roots within a few ulp, fused multiply-adds, no work that only reproduces
the exact arithmetic's roundings). The line reports the fast mode, with its
per-field relative L1 and Linf against the exact mode's state after the same steps (the accuracy
bar is 1e-10) and the exact mode's own figures beside it; --arith exact reports the exact mode
alone.

Prints one JSON line: cell-updates/s, plus the HBM roofline of the dominant sweep kernel
(algorithmic bytes = one FP64 read + write of the state per interior cell = 64 B) timed with HIP
events on the library's own stream, and the CPU baseline (the oracle on the host cores, a
bounded prefix of the same workload).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "cfd-art_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
_T0 = time.time()


def log(msg):
    """Progress on stderr (the JSON result line is the only stdout output)."""
    print(f"[bench {time.time() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def read_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, cfl, steps, steps_1t, T=math.inf, label="euler_2048"):
    """The oracle on the host's cores over the first `steps` steps of a workload (to T when T is
    finite), and on one thread over its first `steps_1t` steps (a scalar reference point)."""
    import cfdart
    import oracle
    d, h = cfg()
    cells = d.nx * d.ny
    L = oracle.lib()
    cores = L.orc_get_threads()
    sim = oracle.Sim(d, h)
    t0 = time.perf_counter()
    k, _ = sim.run(cfdart.ALTERNATING, steps, T=T, cfl=cfl)
    el = time.perf_counter() - t0
    span = f"all {k} steps to t = {T}" if math.isfinite(T) else f"first {k} steps"
    out = {
        "value": cells * k / el,
        "unit": "cell-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle (FP64 C restatement of the reference), {span} of {label} from its initial condition, "
                  f"{el:.1f} s on {cores} OpenMP threads ({cpu_model()})",
    }
    if steps_1t > 0:
        prev = L.orc_set_threads(1)
        try:
            sim1 = oracle.Sim(d, h)
            t0 = time.perf_counter()
            k1, _ = sim1.run(cfdart.ALTERNATING, steps_1t, T=T, cfl=cfl)
            el1 = time.perf_counter() - t0
        finally:
            L.orc_set_threads(prev)
        out["one_thread"] = {"value": cells * k1 / el1, "unit": "cell-updates/s", "cores": 1,
                             "sample": f"first {k1} steps, {el1:.1f} s"}
    return out


def bench_batched(args):
    """Batched workload: all grids advance together (per sweep phase one x and one y launch per
    model over the tiles of every grid, a captured graph per step) versus the same grids one at a
    time with the plain solver."""
    import numpy as np
    import cfdart

    cases = cfdart.configs.batched_mix(args.grids)
    mode = cfdart.ARITH_EXACT if args.arith == "exact" else cfdart.ARITH_FAST
    table = cfdart.autotune.tune(classes=(128, 256, 512, 1024), log=log, arith=mode)
    steps = args.batch_steps
    cells = [c[0].nx * c[0].ny for c in cases]
    work = float(sum(cells)) * steps
    log(f"batched: {len(cases)} grids, {sum(cells)/1e6:.1f} M cells, {steps} steps each, "
        f"{sum(1 for c in cases if c[0].model == 0)} Euler")

    # Tile shapes for the batch: tuned per model on a batch of that model's grids alone (the
    # per-grid table above is tuned for a grid that has the GPU to itself; in a batch the
    # per-tile cost of the work-list favours larger tiles).
    batch_shape = {}
    for model in (cfdart.EULER, cfdart.ALLAIRE):
        sub = [c for c in cases if c[0].model == model]
        if not sub:
            continue
        tb = cfdart.Batch(len(sub))
        tb.set_arith(mode)  # tuned in the arithmetic the batch runs
        ts = [tb.add(d, h, splitting=cfdart.ALTERNATING, max_steps=steps, cfl=cfl) for d, h, cfl in sub]
        batch_shape.update(cfdart.autotune.tune_batch(tb, {model: ts}, steps=3, log=log))
        tb.close()

    b = cfdart.Batch(len(cases))
    b.set_arith(mode)
    slots = []
    for d, h, cfl in cases:
        yh, xr = batch_shape[d.model]
        slots.append(b.add(d, h, splitting=cfdart.ALTERNATING, max_steps=steps, cfl=cfl, y_strip_rows=yh,
                           x_rows_per_wave=xr))
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    log(f"batched: {t_batch:.3f} s -> {work / t_batch / 1e9:.2f} G cell-updates/s")

    t_plain = 0.0
    same = 0
    for (d, h, cfl), s in zip(cases, slots):
        g = cfdart.Grid(d)
        yh, xr = cfdart.autotune.shape_for(table, d.model, d.nx, d.ny)
        g.set_tile(yh, xr)
        g.set_arith(mode)
        g.upload(h)
        g.sync()
        t0 = time.perf_counter()
        g.run(cfdart.ALTERNATING, steps, cfl=cfl)
        t_plain += time.perf_counter() - t0
        a = g.download()[2:-2, 2:-2]
        bb = b.download(s)[2:-2, 2:-2]
        same += int(np.array_equal(a.view(np.uint64), bb.view(np.uint64)))
        g.close()
    log(f"one at a time: {t_plain:.3f} s -> {work / t_plain / 1e9:.2f} G cell-updates/s; bitwise equal {same}/{len(cases)}")
    return {
        "metric": "aggregate cell-updates/s, batched vs one-at-a-time (FP64)",
        "value": work / t_batch, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded disc/bubble generators)",
        "config": {"workload": "batched", "grids": len(cases), "cells": sum(cells),
                   "euler": sum(1 for c in cases if c[0].model == 0), "sides": "log-uniform [128,1024]",
                   "splitting": "alternating", "arith": args.arith},
        "one_at_a_time": {"value": work / t_plain, "seconds": t_plain},
        "speedup_vs_one_at_a_time": t_plain / t_batch,
        "bitwise_equal_alone": f"{same}/{len(cases)}",
        "autotune": {str(m): {str(c): v for c, v in t.items()} for m, t in table.items()},
        "batch_tile": {("euler" if m == 0 else "allaire"): list(v) for m, v in batch_shape.items()},
    }


def bench_served(args):
    """Served stream: Poisson arrivals (50/s, 70 % Euler, sides log-uniform in [64, 2048], steps
    uniform in [50, 2000], a frame every 50 steps, 200 GB admission budget, at most 32 M live
    cells) replayed in real time through the continuous-batching server, then the same requests
    one at a time with the plain solver. The default window (10 s) keeps the run within a minute or so;
    --serve-window 120 replays the configured stream. The offered load (about 31 G cell-updates/s)
    exceeds what one GPU serves, so requests queue. The server prints a progress line every 30 s."""
    import cfdart
    from cfdart import serve

    reqs = serve.trace(seed=7, rate=50.0, duration=args.serve_window)
    work = float(sum(q.n * q.n * q.steps for q in reqs))
    log(f"served: {len(reqs)} requests over {args.serve_window:.1f} s, {work / 1e9:.1f} G cell-updates offered "
        f"({work / args.serve_window / 1e9:.1f} G/s)")
    tile = {0: (cfdart.Y_TRANSPOSED, 8), 1: (cfdart.Y_TRANSPOSED, 8)}
    plain = {0: (cfdart.Y_TRANSPOSED, 4), 1: (cfdart.Y_TRANSPOSED_NARROW, 2)}
    # the server admits while at most 32 M cells are live (enough to fill the GPU); the rest wait
    # first come first served. With --serve-no-cap the same trace without the cap is reported beside it.
    live_cap = 32 * 10**6

    def opts(cap):
        return serve.opts(mem_budget=200 * 10**9, capacity=8192, chunk_steps=10, frame_every=50,
                          frame_host_bytes=2 << 30, realtime=True, tile=tile, plain_tile=plain, live_cells=cap,
                          arith=cfdart.ARITH_EXACT if args.arith == "exact" else cfdart.ARITH_FAST,
                          progress_every=30.0)

    o = opts(live_cap)
    rb, sb = serve.serve_batched(reqs, o)
    lb = serve.latency_stats(reqs, rb)
    log(f"served batched, at most {live_cap / 1e6:.0f} M live cells: makespan {sb.makespan:.2f} s, "
        f"{sb.cell_updates / sb.makespan / 1e9:.2f} G cell-updates/s, p50 {lb['p50']:.2f} s p99 {lb['p99']:.2f} s, "
        f"max live {sb.max_live}, frames {sb.frames} ({sb.frame_bytes / 1e9:.1f} GB)")
    ru = su = lu = None
    if args.serve_no_cap:
        ru, su = serve.serve_batched(reqs, opts(0))
        lu = serve.latency_stats(reqs, ru)
        log(f"served batched, no cap: makespan {su.makespan:.2f} s, {su.cell_updates / su.makespan / 1e9:.2f} G "
            f"cell-updates/s, p50 {lu['p50']:.2f} s p99 {lu['p99']:.2f} s, max live {su.max_live}")
    ro, so = serve.serve_one_at_a_time(reqs, o)
    lo = serve.latency_stats(reqs, ro)
    busy = sum(r.service for r in ro)
    log(f"one at a time: makespan {so.makespan:.2f} s (busy {busy:.2f} s), "
        f"{so.cell_updates / so.makespan / 1e9:.2f} G cell-updates/s, p50 {lo['p50']:.2f} s p99 {lo['p99']:.2f} s")

    def same(ra, rb_):
        return f"{sum(1 for a, b in zip(ra, rb_) if a.status == 0 and a.checksum == b.checksum)}/{len(reqs)}"

    return {
        "metric": "served aggregate cell-updates/s (FP64), continuous batching vs one-at-a-time",
        "value": sb.cell_updates / sb.makespan, "unit": "cell-updates/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded Poisson trace, seeded disc/bubble ICs)",
        "config": {"workload": "served", "requests": len(reqs), "rate_per_s": 50.0, "window_s": args.serve_window,
                   "euler_frac": 0.7, "sides": "log-uniform [64,2048]", "steps": "uniform [50,2000]",
                   "frame_every": 50, "mem_budget_bytes": 200 * 10**9, "live_cells": live_cap, "chunk_steps": 10,
                   "tile": {"euler": tile[0], "allaire": tile[1]}, "arith": args.arith},
        "latency_s": {"batched": lb, "one_at_a_time": lo},
        "one_at_a_time": {"value": so.cell_updates / so.makespan, "makespan_s": so.makespan, "busy_s": busy},
        "makespan_s": sb.makespan,
        "speedup_vs_one_at_a_time": (sb.cell_updates / sb.makespan) / (so.cell_updates / so.makespan),
        "max_live": sb.max_live, "peak_state_bytes": sb.peak_state_bytes,
        "frames": sb.frames, "frame_bytes": sb.frame_bytes,
        "bitwise_equal_alone": same(rb, ro),
        "no_live_cap": None if su is None else {
            "value": su.cell_updates / su.makespan, "makespan_s": su.makespan, "latency_s": lu,
            "max_live": su.max_live, "bitwise_equal_alone": same(ru, ro)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=160,
                    help="oracle steps of euler_2048 for the CPU baseline, about 10 s on 16 threads (0: skip)")
    ap.add_argument("--cpu-steps-1t", type=int, default=8,
                    help="oracle steps of euler_2048 on one thread, about 10 s (0: skip)")
    ap.add_argument("--strip-rows", type=int, default=0)
    ap.add_argument("--tile", default="", help="Y_STRIP_ROWS,X_ROWS_PER_WAVE: fixed tile shape, no autotuning")
    ap.add_argument("--arith", default="fast", choices=["fast", "exact", "fast-only"],
                    help="single-grid workloads: fast = time the exact and the fast arithmetic and report the fast "
                         "one with its accuracy against the exact state; exact = the exact arithmetic only; "
                         "fast-only = the fast arithmetic alone (for profiling). Batched "
                         "and served workloads run the batch and its one-at-a-time baseline in the given arithmetic")
    ap.add_argument("--workload", default="euler_2048",
                    choices=["euler_2048", "allaire_1024", "euler_200", "euler_slab", "godunov_2048", "batched",
                             "served"])
    ap.add_argument("--serve-window", type=float, default=10.0, help="seconds of the arrival trace to replay")
    ap.add_argument("--serve-no-cap", action="store_true",
                    help="served workload: also replay the trace through the server without the live-cell cap")
    ap.add_argument("--grids", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of the state size (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the workloads run on one device; sharding across GPUs is not "
              "implemented); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import cfdart

    if args.copy_probe:
        # the library's plain copy kernel over the 2048^2 Euler state (4 planes of 2048 x 2048)
        d, h = cfdart.configs.euler_2048()
        g = cfdart.Grid(d)
        g.upload(h)
        _check = cfdart._lib._check
        _check(cfdart.lib().cfdart_copy_state(g.h, 5))
        n = 4 * d.ny * ((d.nx + 31) // 32 * 32)
        g.close()
        print(json.dumps({"copy_probe_kernel": "copy_kernel", "copy_probe_bytes_read": n * 8,
                          "copy_probe_bytes_written": n * 8, "launches": 5}))
        return

    if cfdart.device_count() < 1:
        raise SystemExit("bench.py: no GPU visible")
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "served":
        print(json.dumps(bench_served(args)), flush=True)
        return
    single = {
        "euler_2048": dict(cfg=cfdart.configs.euler_2048, model=cfdart.EULER, cfl=0.8, nc=4, tag="EULER",
                           metric="cell-updates/s (2048^2 FP64 Euler, MUSCL-Hancock + HLLC)",
                           data="synthetic (seeded disc generator)",
                           config={"discs": 64, "bc": "reflective L/R, transmissive B/T", "limiter": "van Leer",
                                   "riemann": "HLLC"},
                           cpu=dict(steps=args.cpu_steps, steps_1t=args.cpu_steps_1t)),
        "euler_200": dict(cfg=cfdart.configs.euler_200, model=cfdart.EULER, cfl=0.8, nc=4, tag="EULER", T=0.2,
                          metric="cell-updates/s (200^2 FP64 Euler to t = 0.2, MUSCL-Hancock + HLLC; the "
                                 "CPU-runnable case)",
                          data="synthetic (seeded disc generator)",
                          config={"discs": 8, "bc": "transmissive", "limiter": "van Leer", "riemann": "HLLC",
                                  "T": 0.2},
                          cpu=dict(steps=100000, steps_1t=100000)),
        "godunov_2048": dict(cfg=cfdart.configs.euler_2048_godunov, model=cfdart.EULER, cfl=0.8, nc=4, tag="EULER",
                             metric="cell-updates/s (2048^2 FP64 Euler, first-order Godunov + HLLC)",
                             data="synthetic (seeded disc generator)",
                             config={"discs": 64, "bc": "reflective L/R, transmissive B/T", "riemann": "HLLC",
                                     "scheme": "Godunov (first order), y strips"}),
        "euler_slab": dict(cfg=cfdart.configs.euler_slab, model=cfdart.EULER, cfl=0.8, nc=4, tag="EULER",
                           metric="cell-updates/s per GPU (2048x256 y-slab of 2048^2 FP64 Euler: one GPU's share of 8)",
                           data="synthetic (rows 768-1023 of the seeded 2048^2 disc state)",
                           config={"slab": "2048x256 of 2048x2048", "bc": "reflective L/R, transmissive B/T "
                                   "(stand-in for halo rows)", "limiter": "van Leer", "riemann": "HLLC"}),
        "allaire_1024": dict(cfg=cfdart.configs.allaire_1024, model=cfdart.ALLAIRE, cfl=0.5, nc=6, tag="ALLAIRE",
                             metric="cell-updates/s (1024^2 FP64 Allaire five-equation, MUSCL-Hancock + PVRS-HLLC)",
                             data="synthetic (seeded bubble generator, planar shock)",
                             config={"bubbles": 16, "bc": "transmissive", "limiter": "minmod",
                                     "riemann": "PVRS-HLLC", "z_clip": [1e-6, 1 - 1e-6]},
                             cpu=dict(steps=100, steps_1t=4)),
    }[args.workload]
    T_end = single.get("T", math.inf)
    if math.isfinite(T_end):  # a run to T: every step of it is timed, no warm-up steps
        args.warmup = 0
    d, h = single["cfg"]()
    cfl = single["cfl"]
    cells = d.nx * d.ny
    modes = {"fast": ["exact", "fast"], "exact": ["exact"], "fast-only": ["fast"]}[args.arith]
    arith_of = {"exact": cfdart.ARITH_EXACT, "fast": cfdart.ARITH_FAST}
    # each arithmetic runs at the tile shape tuned in that arithmetic
    tiles = {}
    for m in modes:
        if args.tile:
            yh, xr = (int(v) for v in args.tile.split(","))
            tiles[m] = {"y_strip_rows": yh, "x_rows_per_wave": xr}
        else:
            tiles[m] = cfdart.autotune.tune_one(single["model"], d.nx, log=log, case=(d, h), arith=arith_of[m])
    bytes_per_launch = 16.0 * single["nc"] * cells  # one FP64 read + write of the state per cell

    def timed(mode):
        """The workload's timed run in one arithmetic mode: warm-up steps, clocks brought up on a
        scratch copy, the per-kernel event averages from a pass on the scratch copy, then the
        timed steps without events."""
        arith = arith_of[mode]
        tuned = tiles[mode]
        g = cfdart.Grid(d)
        g.set_tile(tuned["y_strip_rows"], tuned["x_rows_per_wave"])
        g.set_arith(arith)
        if args.strip_rows:
            g.set_strip_rows(args.strip_rows)
        g.upload(h)
        g.run(cfdart.ALTERNATING, args.warmup, cfl=cfl)
        # bring the GPU to its working clocks on a scratch copy of the same state, so that the
        # timed steps of the measured grid (steps warmup .. warmup+steps of the run) are not taken
        # at idle clocks; the measured grid itself is not advanced by this
        scratch = cfdart.Grid(d)
        scratch.set_tile(tuned["y_strip_rows"], tuned["x_rows_per_wave"])
        scratch.set_arith(arith)
        scratch.upload(h)
        steps = args.steps
        if math.isfinite(T_end):
            # the scratch run to T also gives the number of steps the timed run to T takes
            t_s, n_s = scratch.run(cfdart.ALTERNATING, 20000, T=T_end, cfl=cfl)
            if t_s != T_end:
                raise SystemExit(f"bench.py: {args.workload} did not reach t = {T_end} in 20000 steps")
            steps = n_s
            log(f"{args.workload}: the run to t = {T_end} takes {n_s} steps, all timed")
        else:
            scratch.run(cfdart.ALTERNATING, 100, cfl=cfl)
        # the timed steps run without events; the per-kernel averages come from a second pass of
        # the same length on the scratch grid, with an event pair around every sweep (the events
        # add a few microseconds per step, which the timed steps must not carry)
        scratch.set_profiling(True)
        scratch.run(cfdart.ALTERNATING, steps, T=math.inf, cfl=cfl, sync=False)
        prof = scratch.profile()
        scratch.close()
        g.sync()
        t0 = time.perf_counter()
        g.run(cfdart.ALTERNATING, steps, T=T_end, cfl=cfl, sync=False)
        g.sync()
        el = time.perf_counter() - t0
        t_sim, n_sim = g.run(cfdart.ALTERNATING, 0, T=T_end, cfl=cfl)
        per_kernel = {}
        for axis, (ms, nl) in prof.items():
            if nl:
                per_kernel[axis] = {"avg_us": 1e3 * ms / nl, "launches": nl,
                                    "gbs": bytes_per_launch / (ms / nl * 1e-3) / 1e9}
        state = g.download()
        r = dict(value=cells * steps / el, steps=steps, el=el, kernels=per_kernel, counters=g.stats(),
                 t_sim=t_sim, n_sim=n_sim, state=state)
        g.close()
        return r

    if not math.isfinite(T_end):
        log(f"{args.workload}: warm-up {args.warmup} steps, timed {args.steps} steps")
    res = {m: timed(m) for m in modes}
    for m in modes:
        log(f"{args.workload}: {m} arithmetic {res[m]['value'] / 1e9:.2f} G cell-updates/s, "
            f"x {res[m]['kernels']['x']['avg_us']:.1f} us, y {res[m]['kernels']['y']['avg_us']:.1f} us")
    head = res[modes[-1]]
    per_kernel = head["kernels"]
    dom = max(per_kernel, key=lambda a: per_kernel[a]["avg_us"] * per_kernel[a]["launches"])
    tuned = tiles[modes[-1]]
    transposed = (dom == "y" and tuned["y_strip_rows"] < 0 and not args.strip_rows
                  and d.flux_solver == cfdart.MUSCL_HANCOCK)  # the Godunov scheme runs y strips
    kname = f"sweep_{'yt' if transposed else dom}_kernel<{single['tag']}>"
    achieved = per_kernel[dom]["gbs"]
    traffic = read_traffic(args.workload, dom)

    out = {
        "metric": single["metric"],
        "value": head["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["el"] / head["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": single["data"],
        "config": dict({"workload": args.workload, "nx": d.nx, "ny": d.ny}, **single["config"],
                       splitting="alternating", cfl=cfl, sweeps_per_step=2, t_end=head["t_sim"],
                       steps_total=head["n_sim"], arith=modes[-1]),
        "roofline": {"bound": "hbm", "kernel": kname + (" (fast arithmetic)" if modes[-1] == "fast" else ""),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes_per_launch": bytes_per_launch},
        "kernels": per_kernel,
        "tile": {"y_strip_rows": tuned["y_strip_rows"], "x_rows_per_wave": tuned["x_rows_per_wave"]},
        "fallback_counters": head["counters"],
    }
    if modes[-1] == "fast" and "exact" in res:
        ex = res["exact"]
        ex_dom = ex["kernels"][dom]["gbs"]
        out["exact_arith"] = {"value": ex["value"], "ms_per_step": 1e3 * ex["el"] / ex["steps"],
                              "tile": {"y_strip_rows": tiles["exact"]["y_strip_rows"],
                                       "x_rows_per_wave": tiles["exact"]["x_rows_per_wave"]},
                              "kernels": ex["kernels"], "roofline_frac": ex_dom / HBM_PEAK_GBS,
                              "note": "every division and square root correctly rounded: bitwise the "
                                      "reference's arithmetic (the oracle), the library's default"}
        a, b = head["state"], ex["state"]
        ng = d.ng
        errs = []
        if head["n_sim"] == ex["n_sim"]:
            for k in range(a.shape[-1]):
                x, y = a[ng:-ng, ng:-ng, k], b[ng:-ng, ng:-ng, k]
                errs.append((float(np.abs(x - y).sum() / np.abs(y).sum()), float(np.abs(x - y).max() / np.abs(y).max())))
        out["accuracy"] = {
            "against": "the exact mode's state after the same steps (bitwise the CPU reference's arithmetic)",
            "steps": head["n_sim"], "bar": 1e-10,
            "rel_l1_per_field": [e[0] for e in errs], "rel_linf_per_field": [e[1] for e in errs],
            "rel_l1_max": max(e[0] for e in errs) if errs else None,
            "rel_linf_max": max(e[1] for e in errs) if errs else None,
            "t_end_exact": ex["t_sim"]}
        if errs:
            log(f"{args.workload}: fast vs exact after {head['n_sim']} steps: rel L1 max "
                f"{out['accuracy']['rel_l1_max']:.2e}, rel Linf max {out['accuracy']['rel_linf_max']:.2e}")
        else:
            # a run to T whose dt sequences differ by rounding can end on different step counts
            out["accuracy"]["step_mismatch"] = {"fast": head["n_sim"], "exact": ex["n_sim"]}
            log(f"{args.workload}: fast and exact runs ended on different step counts "
                f"({head['n_sim']} vs {ex['n_sim']}); no field comparison")
    if args.cpu_steps > 0 and single.get("cpu"):
        c = single["cpu"]
        out["cpu_baseline"] = cpu_baseline(single["cfg"], cfl, c["steps"], c["steps_1t"], T=single.get("T", math.inf),
                                           label=args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
