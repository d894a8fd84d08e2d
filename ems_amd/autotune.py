"""Autotuner for the batched FP32 form's launch shapes.

The batched step runs one wave per (block of 64 columns, tracer).  Two
shape parameters change its speed but not its results:
  * the block shape, tile_x x 64 / tile_x columns (tile_x = 8, 16, 32 or
    64): a wave walks from the deepest bed of its block to the top, so
    blocks over uneven bathymetry waste lanes on the shallow columns
    (wide blocks bin columns of one row together, square blocks columns
    of one patch), while wider blocks read longer aligned row segments;
  * the strip width (tile_y, columns): blocks are ordered by vertical
    strips, row by row inside a strip, and the sequence is cut into 8
    work-balanced chunks, one per XCD; the strip width sets how far apart
    a block and the block holding its south halo rows run on one XCD
    (L2 reuse) against how many halo columns fall between XCDs;
  * the depth bins (grid): the blocks of each XCD's chunk are ordered by
    the layers their waves walk, in that many bins, deepest first (1: the
    strip order), so long walks start first and short ones fill the end
    of the chunk, at the cost of the strip order's halo reuse.
Which is fastest depends on the window's shape and depth distribution,
so they are picked by timing a few steps of each candidate with HIP
events on the context's own stream.  Results do not depend on the shape
(the stochastic rounding is keyed by cell, tracer and step, not by the
thread), so tuning never changes the answer, only the time.

    tune_context(ctx)            block and strip of one batched context
    tune_ensemble(ctxs)          one block width for every window of an
                                 ensemble graph
"""
import ctypes as C

from . import _lib
from . import tracer as T

BLOCKS = [16]
STRIPS = [256, 64, 128, 512, 1024]
BINS = [1, 2, 4, 8]


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
    """(block width, strip width, depth bins, workgroups per tracer group)
    of a batched context."""
    v = [C.c_int() for _ in range(4)]
    T._chk(_lib.lib().emstr_launch_shape(ctx.ptr, *[C.byref(x) for x in v]),
           "emstr_launch_shape")
    return tuple(x.value for x in v)


def tune_context(ctx, steps=4, blocks=BLOCKS, strips=STRIPS, bins=BINS):
    """Times `steps` steps (an even number: one graph replay holds two) of
    every block width in `blocks` at the first strip width and bin count,
    then the best block at every strip width, then the best of those at
    every bin count; applies the fastest and returns (best, trials) where
    best is (block width, strip width, depth bins, ms_per_step)."""
    if ctx.form != T.BATCH_F32:
        raise ValueError("tune_context: batched contexts only")
    steps = max(2, steps // 2 * 2)
    L = _lib.lib()
    ev = _Events()
    trials = []

    def run(tx, sw, nb):
        T._chk(L.emstr_tune(ctx.ptr, tx, sw, nb, 1), "emstr_tune")
        ctx.step(2)  # captures the graph for this shape
        ctx.sync()
        ms = ev.time(ctx.stream, lambda: ctx.step(steps)) / steps
        trials.append((tx, sw, nb, ms))
        return ms

    for tx in blocks:
        run(tx, strips[0], bins[0])
    tx = min(trials, key=lambda r: r[3])[0]
    for sw in strips[1:]:
        run(tx, sw, bins[0])
    sw = min(trials, key=lambda r: r[3])[1]
    for nb in bins[1:]:
        run(tx, sw, nb)
    best = min(trials, key=lambda r: r[3])
    T._chk(L.emstr_tune(ctx.ptr, best[0], best[1], best[2], 1), "emstr_tune")
    ctx.step(2)
    ctx.sync()
    ev.close()
    return best, trials


def tune_ensemble(ctxs, steps=4, blocks=BLOCKS):
    """One block width for every window of an ensemble graph: each width
    is timed on a fresh ensemble graph and the fastest is kept.  Returns
    (best_width, ms_per_step, trials) and leaves the contexts tuned to it."""
    steps = max(2, steps // 2 * 2)
    L = _lib.lib()
    ev = _Events()
    trials = []

    def apply(tx):
        for c in ctxs:
            T._chk(L.emstr_tune(c.ptr, tx, 0, 0, 1), "emstr_tune")

    for tx in blocks:
        apply(tx)
        ens = T.Ensemble(ctxs)
        ens.step(2)
        ens.sync()
        ms = ev.time(ens.stream, lambda: ens.step(steps)) / steps
        ens.sync()
        ens.close()
        trials.append((tx, ms))
    best = min(trials, key=lambda r: r[1])
    apply(best[0])
    ev.close()
    return best[0], best[1], trials
