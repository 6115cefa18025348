"""Benchmark: FP64 2D Euler MUSCL-Hancock on one MI355X, 2048x2048.

Workload "euler_2048": 64 seeded high-pressure discs (radii U[0.005,0.03]) on an ambient gas,
reflective left/right, transmissive bottom/top, van Leer + HLLC, alternating x/y splitting,
CFL 0.8 (the first 5 steps at 0.2, as the reference's compute_dt does). A step is one full split
time step (an x and a y sweep + the CFL reduction) on the device-resident grid; the state is
uploaded before the timed region, dt never leaves the device.

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

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "cfd-art_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def read_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_euler_2048.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def cpu_baseline(steps):
    import cfdart
    import oracle
    d, h = cfdart.configs.euler_2048()
    sim = oracle.Sim(d, h)
    t0 = time.perf_counter()
    k, _ = sim.run(cfdart.ALTERNATING, steps, cfl=0.8)
    el = time.perf_counter() - t0
    cells = d.nx * d.ny
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": cells * k / el,
        "unit": "cell-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle (FP64 C restatement of the reference), first {k} steps of euler_2048 "
                  f"from its initial condition, {el:.1f} s on {cores} OpenMP threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=12, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--strip-rows", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of the state size (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the workload is a single 2048^2 grid); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import cfdart

    if args.copy_probe:
        import torch
        n = 2048 * 2048 * 4
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 5}))
        return

    if cfdart.device_count() < 1:
        raise SystemExit("bench.py: no GPU visible")
    d, h = cfdart.configs.euler_2048()
    cells = d.nx * d.ny
    g = cfdart.Grid(d)
    if args.strip_rows:
        g.set_strip_rows(args.strip_rows)
    g.upload(h)
    g.run(cfdart.ALTERNATING, args.warmup, cfl=0.8)
    g.set_profiling(True)
    g.sync()
    t0 = time.perf_counter()
    g.run(cfdart.ALTERNATING, args.steps, cfl=0.8, sync=False)
    g.sync()
    el = time.perf_counter() - t0
    prof = g.profile()
    g.set_profiling(False)
    t_sim, n_sim = g.run(cfdart.ALTERNATING, 0, cfl=0.8)
    counters = g.stats()

    bytes_per_launch = 64.0 * cells  # one FP64 read + write of the 4-component state per cell
    per_kernel = {}
    for axis, (ms, nl) in prof.items():
        if nl:
            per_kernel[axis] = {"avg_us": 1e3 * ms / nl, "launches": nl,
                                "gbs": bytes_per_launch / (ms / nl * 1e-3) / 1e9}
    dom = max(per_kernel, key=lambda a: per_kernel[a]["avg_us"] * per_kernel[a]["launches"])
    kname = {"x": "sweep_x_kernel<EULER>", "y": "sweep_y_kernel<EULER>"}[dom]
    achieved = per_kernel[dom]["gbs"]
    traffic = read_traffic(dom)

    value = cells * args.steps / el
    out = {
        "metric": "cell-updates/s (2048^2 FP64 Euler, MUSCL-Hancock + HLLC)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded disc generator)",
        "config": {"workload": "euler_2048", "nx": d.nx, "ny": d.ny, "discs": 64,
                   "bc": "reflective L/R, transmissive B/T", "limiter": "van Leer",
                   "riemann": "HLLC", "splitting": "alternating", "cfl": 0.8,
                   "sweeps_per_step": 2, "t_end": t_sim, "steps_total": n_sim},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "kernels": per_kernel,
        "fallback_counters": counters,
    }
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
