"""Autotuner for the batched FP32 form's launch shapes.

The batched kernels walk cells and columns in tile order (tile_x x tile_y
columns per tile) over persistent grids sized to fill the GPU.  Which tile
keeps a wave's neighbours in L2 best, and how many workgroups a window
should occupy when several windows share the GPU (an ensemble graph),
depend on the window's shape and depth distribution, so they are picked
by timing a few steps of each candidate with HIP events on the context's
own stream.  Results do not depend on the shape (the stochastic rounding
is keyed by cell, tracer and step, not by the thread), so tuning never
changes the answer, only the time.

    tune_context(ctx)            tiles and grid of one batched context
    tune_ensemble(ctxs)          per-window grids, in proportion to each
                                 window's cells, for one ensemble graph
"""
import ctypes as C

from . import _lib
from . import tracer as T

TILES = [(32, 32), (16, 16), (64, 16), (16, 64), (32, 8), (8, 32), (64, 64)]
GRID_SCALES = [1.0, 0.5, 0.25]


class _Events:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]
        self.a, self.b = C.c_void_p(), C.c_void_p()
        if self.h.hipEventCreate(C.byref(self.a)) or self.h.hipEventCreate(C.byref(self.b)):
            raise RuntimeError("hipEventCreate failed")

    def time(self, stream, fn):
        if self.h.hipEventRecord(self.a, C.c_void_p(stream)):
            raise RuntimeError("hipEventRecord failed")
        fn()
        if self.h.hipEventRecord(self.b, C.c_void_p(stream)) or \
                self.h.hipEventSynchronize(self.b):
            raise RuntimeError("hipEventRecord/Synchronize failed")
        t = C.c_float()
        if self.h.hipEventElapsedTime(C.byref(t), self.a, self.b):
            raise RuntimeError("hipEventElapsedTime failed")
        return float(t.value)

    def close(self):
        self.h.hipEventDestroy(self.a)
        self.h.hipEventDestroy(self.b)


def shape(ctx):
    """(tile_x, tile_y, grid, full_grid) of a batched context."""
    v = [C.c_int() for _ in range(4)]
    T._chk(_lib.lib().emstr_launch_shape(ctx.ptr, *[C.byref(x) for x in v]),
           "emstr_launch_shape")
    return tuple(x.value for x in v)


def tune_context(ctx, steps=4, tiles=TILES, scales=GRID_SCALES):
    """Times `steps` steps (an even number: one graph replay holds two) of
    every tile in `tiles` at the full grid, then the best tile at each
    grid scale; applies the fastest and returns (best, trials) where best
    is (tile_x, tile_y, grid, ms_per_step)."""
    if ctx.form != T.BATCH_F32:
        raise ValueError("tune_context: batched contexts only")
    steps = max(2, steps // 2 * 2)
    L = _lib.lib()
    full_grid = shape(ctx)[3]
    ev = _Events()
    trials = []

    def run(tx, ty, grid):
        T._chk(L.emstr_tune(ctx.ptr, tx, ty, grid, 1), "emstr_tune")
        ctx.step(2)  # captures the graph for this shape
        ctx.sync()
        ms = ev.time(ctx.stream, lambda: ctx.step(steps)) / steps
        trials.append((tx, ty, grid, ms))
        return ms

    for tx, ty in tiles:
        run(tx, ty, full_grid)
    tx, ty, _, _ = min(trials, key=lambda r: r[3])
    for s in scales:
        if s != 1.0:
            run(tx, ty, max(8, int(full_grid * s) // 8 * 8))
    best = min(trials, key=lambda r: r[3])
    T._chk(L.emstr_tune(ctx.ptr, best[0], best[1], best[2], 1), "emstr_tune")
    ctx.step(2)
    ctx.sync()
    ev.close()
    return best, trials


def tune_ensemble(ctxs, steps=4, scales=(1.0, 0.5, 0.25, 0.125, 0.0625)):
    """Per-window advection grids for one ensemble graph: window i gets
    max(8, s * full_grid * cells_i / max_cells) workgroups; every scale s
    is timed on a fresh ensemble graph and the fastest is kept.  Returns
    (best_scale, ms_per_step, trials) and leaves the contexts tuned to it."""
    steps = max(2, steps // 2 * 2)
    L = _lib.lib()
    ev = _Events()
    cells = [c.case.nwet for c in ctxs]
    cmax = float(max(cells))
    full_grid = shape(ctxs[0])[3]
    trials = []

    def apply(s):
        for c, n in zip(ctxs, cells):
            g = max(8, int(s * full_grid * n / cmax) // 8 * 8)
            T._chk(L.emstr_tune(c.ptr, 0, 0, g, 1), "emstr_tune")

    for s in scales:
        apply(s)
        ens = T.Ensemble(ctxs)
        ens.step(2)
        ens.sync()
        ms = ev.time(ens.stream, lambda: ens.step(steps)) / steps
        ens.sync()
        ens.close()
        trials.append((s, ms))
    best = min(trials, key=lambda r: r[1])
    apply(best[0])
    ev.close()
    return best[0], best[1], trials
