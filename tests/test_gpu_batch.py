"""GPU batched form (one fused persistent kernel over all tracers and
columns, FP32 with FP64 budgets / refinement / thicknesses, hipGraph
replay) against the FP64 oracle and the bit-exact plain form.
Tolerances: max |error| <= 1e-5 of each tracer's range, relative mass
drift <= 1e-6."""
import sys

import numpy as np
import pytest

from ems_amd import configs, structs as S
from ems_amd import tracer as T
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_batch(case, nsteps, graph=1, tile=None):
    ctx = T.Context(case, T.BATCH_F32)
    if tile:
        ctx.tune(tile[0], tile[1], 0, graph)
    else:
        ctx.tune(0, 0, 0, graph)
    ctx.upload()
    m0 = ctx.mass()
    ctx.step(nsteps)
    ctx.sync()
    m1 = ctx.mass()
    ctx.download()
    out = case.tracers()
    ctx.close()
    return out, m0, m1


def errors(a, ref):
    rng = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-30)
    return (np.abs(a - ref) / rng).max(axis=1)


def oracle_steps(case, snap, n):
    H.restore(case, snap)
    for _ in range(n):
        assert O.step(case) == 0
    return case.tracers()


def test_c1_batch_vs_oracle_200_steps():
    say("batch c1 200 steps")
    case = configs.c1()
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    b, m0, m1 = run_batch(case, 200)
    ref = oracle_steps(case, snap, 200)
    e = errors(b, ref)
    assert e.max() <= 1e-5, e
    assert np.all(np.abs(m1 - m0) <= 1e-6 * np.abs(m0))


@pytest.mark.parametrize("seed", [41, 42])
def test_small_batch_vs_oracle(seed):
    say("batch small", seed)
    case = configs.small(seed=seed, ntr=5, dt=120.0)
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    b, m0, m1 = run_batch(case, 60)
    ref = oracle_steps(case, snap, 60)
    e = errors(b, ref)
    assert e.max() <= 1e-5, e


def test_batch_graph_matches_direct_launches():
    say("batch graph vs direct")
    case = configs.small(seed=43, ntr=3)
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    a, _, _ = run_batch(case, 7, graph=1)
    H.restore(case, snap)
    b, _, _ = run_batch(case, 7, graph=0)
    assert np.array_equal(a, b)
    H.restore(case, snap)
    c, _, _ = run_batch(case, 7, graph=1, tile=(16, 4))
    assert np.array_equal(a, c)


def test_batch_moving_surface():
    say("batch moving surface")
    case = configs.small(seed=13, ntr=3, nz=6, dmin=2.0, dmax=30.0)
    rng = np.random.default_rng(4)
    g = case.g
    w2 = S.view(g.w2_t, case.sgsizS, np.int32)[1:g.v2_t + 1]
    dl = case.layers[1] - case.layers[0]
    old = case.a2(case.wc.oldeta)
    eta = case.a2(case.wd.eta)
    old[w2] = rng.uniform(-1.3, 0.0, len(w2)) * dl
    eta[w2] = old[w2]
    O.set_dz(case.gp, case.wp, case.pp)
    eta[w2] = old[w2] + rng.uniform(-0.9, 0.9, len(w2)) * dl
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    b, _, _ = run_batch(case, 1)
    ref = oracle_steps(case, snap, 1)
    e = errors(b, ref)
    assert e.max() <= 1e-5, e


def test_batch_stiff_refinement():
    say("batch stiff")
    case = H.box(nxi=6, nyi=5, nz=20, depth=10.0, ntr=2)
    H.scheme(case, S.VANLEER, 0, 600.0)
    nz, ny, nx = case.nz, case.ny, case.nx
    rng = np.random.default_rng(5)
    H.set_dense_face_fields(case, u1=np.zeros((nz, ny, nx)), u2=np.zeros((nz, ny, nx)),
                            w=np.zeros((nz, ny, nx)),
                            Kz=rng.uniform(0.01, 1.0, (nz, ny, nx)))
    for n in range(2):
        H.set_tracer_dense(case, n, rng.uniform(1.0, 2.0, (nz, ny, nx)))
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    b, m0, m1 = run_batch(case, 3)
    ref = oracle_steps(case, snap, 3)
    e = errors(b, ref)
    assert e.max() <= 1e-5, e
    assert np.all(np.abs(m1 - m0) <= 1e-6 * np.abs(m0))


def test_c3_batch_vs_plain_full_size():
    """Full 400x600x40, 40 tracers: the batched FP32 form against the FP64
    plain form (bit-exact with the oracle at small sizes) after 20 steps;
    mass drift of both."""
    say("batch c3 full size vs plain, 20 steps")
    from ems_amd import _lib
    case = configs.c3()
    L = _lib.lib()
    L.set_dz(case.gp, case.wp, case.pp)
    L.emstr_forget_window(case.gp)
    p = T.Context(case, T.PLAIN_F64)
    p.upload_fields()
    p.vel_w()
    p.upload()
    b = T.Context(case, T.BATCH_F32)
    b.upload()
    mp0, mb0 = p.mass(), b.mass()
    p.step(20)
    b.step(20)
    mp1, mb1 = p.mass(), b.mass()
    p.download()
    ref = case.tracers()
    b.download()
    out = case.tracers()
    p.close()
    b.close()
    e = errors(out, ref)
    say("c3 max err per tracer", e.max(), "relative mass drift",
        (np.abs(mb1 - mb0) / np.abs(mb0)).max())
    assert e.max() <= 1e-5, e
    assert np.all(np.abs(mb1 - mb0) <= 1e-6 * np.abs(mb0))
    assert np.all(np.abs(mp1 - mp0) <= 1e-12 * np.abs(mp0))


def pointwise(a, ref, floor=1e-3):
    """max |a - ref| / |ref| per tracer over the cells where |ref| is at
    least floor times the tracer's largest magnitude"""
    big = np.abs(ref) >= floor * np.abs(ref).max(axis=1, keepdims=True)
    rel = np.where(big, np.abs(a - ref) / np.maximum(np.abs(ref), 1e-300), 0.0)
    return rel.max(axis=1)


def test_c3_batch_1000_steps_error_and_drift():
    """Full c3 over 1000 steps against the FP64 plain form: the error
    (|difference| over each tracer's largest magnitude) stays within 1e-5
    at steps 200, 400 and 1000, and the relative mass drift of every
    tracer within 1e-6, against its own start and against the FP64 form's
    mass at step 1000 (FP64 compensated budgets of the 32-bit state).  The
    pointwise relative error, over the cells holding more than 1e-3 of a
    tracer's largest magnitude, stays within 1e-5 too."""
    say("batch c3 1000 steps: error and drift")
    from ems_amd import _lib
    case = configs.c3()
    L = _lib.lib()
    L.set_dz(case.gp, case.wp, case.pp)
    L.emstr_forget_window(case.gp)
    p = T.Context(case, T.PLAIN_F64)
    p.upload_fields()
    p.vel_w()
    p.upload()
    b = T.Context(case, T.BATCH_F32)
    b.upload()
    mb0 = b.mass()
    done = 0
    for upto in (200, 400, 1000):
        p.step(upto - done)
        b.step(upto - done)
        done = upto
        p.download()
        ref = case.tracers()
        b.download()
        got = case.tracers()
        e = errors(got, ref)
        pw = pointwise(got, ref)
        say("c3 step %d: error %.3g (tracer %d), pointwise above 1e-3: %.3g"
            % (upto, e.max(), int(np.argmax(e)), pw.max()))
        assert e.max() <= 1e-5, (upto, e)
        assert pw.max() <= 1e-5, (upto, pw)
    mb1, mp1 = b.mass(), p.mass()
    p.close()
    b.close()
    drift = np.abs(mb1 - mb0) / np.abs(mb0)
    vs64 = np.abs(mb1 - mp1) / np.abs(mp1)
    say("drift after 1000 steps %.3g, against the FP64 form %.3g" % (drift.max(), vs64.max()))
    assert drift.max() <= 1e-6, drift
    assert vs64.max() <= 1e-6, vs64
