"""Tile-shape autotuner for the sweep kernels.

For each model and grid-size class it times the candidate tile shapes on the device
(cfdart_time_tile: kernel time of one x sweep and one y sweep, HIP events, after warm-up) and
keeps the fastest x-sweep rows-per-wavefront and y-sweep tile (a strip height, or one of the
transposed tiles L.Y_TRANSPOSED and L.Y_TRANSPOSED_NARROW). The x and y choices are independent
(separate kernels). Results are cached as JSON keyed by device name.
"""
import json
import math
import os

from . import _lib as L

X_ROWS = (1, 2, 4, 8, 16)
Y_STRIPS = (8, 16, 32, 64, 128)
SIZE_CLASSES = (128, 256, 512, 1024, 2048)


def size_class(nx, ny):
    side = math.sqrt(nx * ny)
    return min(SIZE_CLASSES, key=lambda c: abs(math.log(side / c)))


def _case(model, n):
    """A representative state of the size class: the workload's own generators (shocks and
    interfaces make the kernels' branches diverge as in production)."""
    from . import configs
    if model == L.EULER:
        return configs.euler_discs(n, 8, (0.03, 0.10), (0, 0, 0, 0), seed=7)
    return configs.allaire_bubbles(n, 16, seed=7)


def _around(best, lo, step):
    """Refinement candidates between best/2 and 2*best (the coarse grid's neighbours)."""
    return sorted({max(lo, int(round(best * f / step)) * step)
                   for f in (0.5, 0.625, 0.75, 0.875, 1.0, 1.125, 1.25, 1.5, 1.75, 2.0)})


def tune_one(model, n, reps=5, log=None, case=None, arith=L.ARITH_EXACT):
    """Coarse search over powers of two, then a finer search around the best of each axis: the
    number of wavefronts a shape launches against the chip's resident-wave slots (256 CUs x
    4 SIMDs x waves per SIMD) decides how full the last round of waves is, which powers of two
    alone do not resolve. The shapes are timed in the arithmetic the grid will run."""
    d, h = case if case is not None else _case(model, n)
    # In the fast arithmetic the MUSCL-Hancock y sweep keeps to the transposed tiles: the strip
    # tile's faces are compiled apart from theirs under FMA contraction, so the two agree to
    # rounding rather than bit for bit, and a grid alone then ends on the bits of the same grid in
    # a batch (whose y sweep is the transposed tile) only on the same tile kind.
    strips = not (arith == L.ARITH_FAST and d.flux_solver == L.MUSCL_HANCOCK)
    tx = {xr: L.time_tile(d, 32, xr, reps, h, arith)[0] for xr in X_ROWS}
    ty = {yh: L.time_tile(d, yh, 4, reps, h, arith)[1] for yh in Y_STRIPS} if strips else {}
    for xr in _around(min(tx, key=tx.get), 1, 1):
        if xr not in tx:
            tx[xr] = L.time_tile(d, 32, xr, reps, h, arith)[0]
    for yh in (_around(min(ty, key=ty.get), 4, 2) if strips else ()):
        if yh not in ty:
            ty[yh] = L.time_tile(d, yh, 4, reps, h, arith)[1]
    for yt in (L.Y_TRANSPOSED, L.Y_TRANSPOSED_NARROW):
        ty[yt] = L.time_tile(d, yt, 4, reps, h, arith)[1]
    bx = min(tx, key=tx.get)
    by = min(ty, key=ty.get)
    if log:
        log(f"autotune model={model} n={n} arith={arith}: x {{{', '.join(f'{k}:{v:.1f}us' for k, v in sorted(tx.items()))}}} -> {bx}; "
            f"y {{{', '.join(f'{k}:{v:.1f}us' for k, v in sorted(ty.items()))}}} -> {by}")
    return {"y_strip_rows": by, "x_rows_per_wave": bx, "x_us": tx[bx], "y_us": ty[by]}


def tune(models=(L.EULER, L.ALLAIRE), classes=SIZE_CLASSES, reps=5, cache=None, log=None, arith=L.ARITH_EXACT):
    """Returns {model: {size_class: {...}}}; reads/writes the JSON cache when given."""
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        return {int(m): {int(c): v for c, v in t.items()} for m, t in table.items()}
    table = {m: {c: tune_one(m, c, reps, log, arith=arith) for c in classes} for m in models}
    if cache:
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(table, f, indent=1)
    return table


def shape_for(table, model, nx, ny):
    e = table[model][size_class(nx, ny)]
    return e["y_strip_rows"], e["x_rows_per_wave"]


def tune_batch(batch, slots_by_model, steps=4, log=None, y_candidates=(L.Y_TRANSPOSED,),
               x_candidates=(2, 4, 8, 16)):
    """Tile shapes for a live batch, per model: coordinate descent (x rows per wavefront, then y
    tile) on the wall time of `steps` whole batch steps. The batch state advances during
    tuning, so callers tune on a copy or before the timed work. Returns {model: (yH, xrows)}."""
    import time
    best = {}
    for model, slots in slots_by_model.items():
        if not slots:
            continue
        cur = [y_candidates[0], 8]

        def measure(yh, xr):
            for s in slots:
                batch.set_tile(s, yh, xr)
            batch.step(1)
            batch.sync()
            t0 = time.perf_counter()
            batch.step(steps)
            batch.sync()
            return (time.perf_counter() - t0) / steps

        tx = {xr: measure(cur[0], xr) for xr in x_candidates}
        cur[1] = min(tx, key=tx.get)
        ty = {yh: measure(yh, cur[1]) for yh in y_candidates}
        cur[0] = min(ty, key=ty.get)
        for s in slots:
            batch.set_tile(s, cur[0], cur[1])
        best[model] = tuple(cur)
        if log:
            log(f"batch autotune model={model}: x {{{', '.join(f'{k}:{v*1e3:.2f}ms' for k, v in tx.items())}}}; "
                f"y {{{', '.join(f'{k}:{v*1e3:.2f}ms' for k, v in ty.items())}}} -> {tuple(cur)}")
    return best
