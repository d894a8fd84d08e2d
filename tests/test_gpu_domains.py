"""Many windows at once (the multi-domain set c5 and the QUICKEST basin
c2): every window keeps its own device mirror behind the drop-in entry
point, as the reference's window loop calls tracer_step_3d per window
(tracers.c:142-209).  The plain form must equal the oracle bit for bit in
every window; the FP32 forms must stay within 1e-5 of each tracer's
range, with relative mass drift within 1e-6, on the stiff windows whose
diffusion numbers reach ~1e4."""
import ctypes as C
import sys

import numpy as np
import pytest

from ems_amd import _lib, configs
from ems_amd import structs as S
from ems_amd import tracer as T
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def errors(a, ref):
    rng = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-30)
    return (np.abs(a - ref) / rng).max(axis=1)


def test_c5_windows_interleaved_bit_exact():
    say("c5: 64 windows, drop-in steps interleaved across windows")
    doms = configs.c5(ndom=64)
    L = _lib.lib()
    snaps = []
    for c in doms:
        c.prepare(O.set_dz)
        snaps.append(H.snapshot(c))
    for _ in range(2):
        for c in doms:
            L.tracer_step_3d(c.gp, c.wp, c.pp)
            msg, cell = _lib.last_error()
            assert msg == "", (c.name, msg, cell)
    gpu = [c.tracers() for c in doms]
    for c, snap, g in zip(doms, snaps, gpu):
        H.restore(c, snap)
        for _ in range(2):
            assert O.step(c, 8) == 0
        ref = c.tracers()
        L.emstr_forget_window(c.gp)
        assert np.array_equal(g, ref), (c.name, np.abs(g - ref).max())


def test_tracer_step_window_loop_bit_exact():
    """tracer_step (tracers.c:142-209) over 12 c5 windows in a shuffled
    twin order, with a master callback that records the copies: every
    window equals the oracle bit for bit, the copies come per window
    (TRACERS) in the twin order, then TMASS for every window."""
    say("tracer_step: 12 windows, twin order, master callbacks")
    doms = configs.c5(ndom=12)
    L = _lib.lib()
    n = len(doms)
    for c in doms:
        c.prepare(O.set_dz)
    snaps = [H.snapshot(c) for c in doms]
    G = (C.POINTER(S.Geometry) * (n + 1))(None, *[c.gp for c in doms])
    W = (C.POINTER(S.Window) * (n + 1))(None, *[c.wp for c in doms])
    P = (C.POINTER(S.WinPriv) * (n + 1))(None, *[c.pp for c in doms])
    order = [0] + list(np.random.default_rng(4).permutation(n) + 1)
    twin = (C.c_int * (n + 1))(*order)
    doms[0].wc.twin = C.cast(twin, C.POINTER(C.c_int))
    seen = []
    wins = {C.addressof(c.g): i + 1 for i, c in enumerate(doms)}
    CB = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(S.Geometry), C.POINTER(S.Window), C.c_int)

    def empty(master, window, windat, mode):
        seen.append((wins[C.addressof(window.contents)], mode))

    cb = CB(empty)
    m = S.Master(mode2d=0, crf=0, win_data_empty_3d=C.cast(cb, C.c_void_p))
    for _ in range(2):
        L.tracer_step(C.byref(m), G, W, P, n)
        msg, cell = _lib.last_error()
        assert msg == "" and m.crf == 0, (msg, cell)
    want = ([(w, S.TRACERS) for w in order[1:]] + [(w, S.TMASS) for w in order[1:]]) * 2
    assert seen == want
    gpu = [c.tracers() for c in doms]
    doms[0].wc.twin = None
    for c, snap, g in zip(doms, snaps, gpu):
        H.restore(c, snap)
        for _ in range(2):
            assert O.step(c, 8) == 0
        ref = c.tracers()
        L.emstr_forget_window(c.gp)
        assert np.array_equal(g, ref), (c.name, np.abs(g - ref).max())


@pytest.mark.parametrize("form", [T.BATCH_F32, T.PREC_F32])
def test_c5_stiff_windows_fp32(form):
    say("c5 stiff windows, form", form)
    doms = [c for c in configs.c5(ndom=64) if c.stiff]
    L = _lib.lib()
    worst = 0.0
    for c in doms:
        L.set_dz(c.gp, c.wp, c.pp)
        L.emstr_forget_window(c.gp)
        p = T.Context(c, T.PLAIN_F64)
        p.upload_fields()
        p.vel_w()
        p.upload()
        f = T.Context(c, form)
        f.upload()
        m0 = f.mass()
        p.step(100)
        f.step(100)
        m1 = f.mass()
        p.download()
        ref = c.tracers()
        f.download()
        got = c.tracers()
        p.close()
        f.close()
        e = errors(got, ref)
        drift = np.abs(m1 - m0) / np.abs(m0)
        worst = max(worst, e.max())
        assert e.max() <= 1e-5, (c.name, e.max())
        assert drift.max() <= 1e-6, (c.name, drift.max())
    say("worst error", worst)


def test_c2_quickest_bit_exact():
    say("c2: 200x300x30 QUICKEST, 10 tracers, 3 steps")
    case = configs.c2()
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    L = _lib.lib()
    for _ in range(3):
        L.tracer_step_3d(case.gp, case.wp, case.pp)
        assert _lib.last_error()[0] == ""
    gpu = case.tracers()
    H.restore(case, snap)
    for _ in range(3):
        assert O.step(case, 16) == 0
    ref = case.tracers()
    L.emstr_forget_window(case.gp)
    assert np.array_equal(gpu, ref), np.abs(gpu - ref).max()


def _batched_pair(c):
    """Two batched contexts of one window from the same host state."""
    L = _lib.lib()
    L.set_dz(c.gp, c.wp, c.pp)
    L.emstr_forget_window(c.gp)
    p = T.Context(c, T.PLAIN_F64)
    p.upload_fields()
    p.vel_w()
    p.close()
    a = T.Context(c, T.BATCH_F32)
    a.upload()
    b = T.Context(c, T.BATCH_F32)
    b.upload()
    return a, b


def test_c5_ensemble_graph_equals_separate_steps():
    """All 64 windows (van Leer, the batched form's scheme) in one
    ensemble graph, against each window stepped on its own: the same
    kernels on the same data, so the results agree bit for bit."""
    say("c5: 64 windows, one ensemble graph vs separate steps")
    doms = configs.c5(ndom=64, trasc=S.VANLEER)
    pairs = [_batched_pair(c) for c in doms]
    for a, _ in pairs:
        a.step(4)
    ens = T.Ensemble([b for _, b in pairs])
    ens.step(4)
    ens.sync()
    ens.close()
    for c, (a, b) in zip(doms, pairs):
        a.download()
        ta = c.tracers()
        b.download()
        tb = c.tracers()
        a.close()
        b.close()
        assert np.array_equal(ta, tb), (c.name, np.abs(ta - tb).max())


def test_ensemble_rejects_odd_steps_and_plain_contexts():
    say("ensemble argument checks")
    doms = configs.c5(ndom=2, trasc=S.VANLEER)
    a, b = _batched_pair(doms[0])
    ens = T.Ensemble([a, b])
    with pytest.raises(RuntimeError, match="even"):
        ens.step(3)
    ens.close()
    a.close()
    b.close()
    p = T.Context(doms[1], T.PLAIN_F64)
    p.upload_fields()
    with pytest.raises(RuntimeError, match="batched"):
        T.Ensemble([p])
    p.close()
