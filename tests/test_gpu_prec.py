"""GPU precision form (FP32 tracers, fluxes and sweeps; FP64 budgets,
refinement pass and layer thicknesses) against the FP64 oracle.
Tolerances: max |error| <= 1e-5 of the tracer's range after the stated
steps, relative mass drift <= 1e-6."""
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


def run_prec(case, nsteps):
    ctx = T.Context(case, T.PREC_F32)
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


def test_c1_fp32_vs_oracle_200_steps():
    say("prec c1 200 steps")
    case = configs.c1()
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    f32, m0, m1 = run_prec(case, 200)
    H.restore(case, snap)
    for _ in range(200):
        O.step(case)
    ref = case.tracers()
    e = errors(f32, ref)
    assert e.max() <= 1e-5, e
    assert np.all(np.abs(m1 - m0) <= 1e-6 * np.abs(m0))


@pytest.mark.parametrize("trasc,ult", [(S.VANLEER, 0), (S.QUICKEST, 0), (S.QUICKEST, 1)])
def test_small_fp32_vs_oracle(trasc, ult):
    say("prec small", trasc, ult)
    case = configs.small(seed=41, ntr=4, trasc=trasc, ultimate=ult, dt=120.0)
    case.prepare(O.set_dz)
    snap = H.snapshot(case)
    f32, m0, m1 = run_prec(case, 50)
    H.restore(case, snap)
    for _ in range(50):
        O.step(case)
    ref = case.tracers()
    e = errors(f32, ref)
    assert e.max() <= 1e-5, e


def test_stiff_columns_use_fp64_refinement():
    """D = Kz*dt/dz^2 up to ~1e4 (0.5 m layers, Kz 1 m2/s, dt 600 s)."""
    say("prec stiff vdiff")
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
    f32, m0, m1 = run_prec(case, 3)
    H.restore(case, snap)
    for _ in range(3):
        O.step(case)
    e = errors(f32, case.tracers())
    assert e.max() <= 1e-5, e


def pointwise(a, ref, floor=1e-3):
    """max |a - ref| / |ref| per tracer over the cells where |ref| is at
    least floor times the tracer's largest magnitude"""
    big = np.abs(ref) >= floor * np.abs(ref).max(axis=1, keepdims=True)
    rel = np.where(big, np.abs(a - ref) / np.maximum(np.abs(ref), 1e-300), 0.0)
    return rel.max(axis=1)


def test_c2_precision_vs_plain_500_steps():
    """The 200x300x30 curvilinear case, 10 tracers, QUICKEST, dt = 120 s:
    the FP32 precision form against the FP64 plain form (bit-exact with
    the oracle) over 500 steps.  Error (|difference| over each tracer's
    largest magnitude) within 1e-5 at steps 100 and 500, the pointwise
    relative error over cells above 1e-3 of the largest magnitude reported,
    and the relative mass drift within 1e-6 against the FP64 form's mass at
    step 500."""
    say("prec c2 500 steps vs plain")
    from ems_amd import _lib
    case = configs.c2()
    L = _lib.lib()
    L.set_dz(case.gp, case.wp, case.pp)
    L.emstr_forget_window(case.gp)
    p = T.Context(case, T.PLAIN_F64)
    p.upload_fields()
    p.vel_w()
    p.upload()
    q = T.Context(case, T.PREC_F32)
    q.upload()
    done = 0
    for upto in (100, 500):
        p.step(upto - done)
        q.step(upto - done)
        done = upto
        p.download()
        ref = case.tracers()
        q.download()
        got = case.tracers()
        e = errors(got, ref)
        say("c2 step %d: error %.3g (tracer %d), pointwise above 1e-3: %.3g"
            % (upto, e.max(), int(np.argmax(e)), pointwise(got, ref).max()))
        assert e.max() <= 1e-5, (upto, e)
    mq, mp = q.mass(), p.mass()
    p.close()
    q.close()
    vs64 = np.abs(mq - mp) / np.abs(mp)
    say("c2 mass against the FP64 form after 500 steps: %.3g" % vs64.max())
    assert vs64.max() <= 1e-6, vs64
