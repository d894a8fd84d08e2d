"""GPU: the FP32 form of the COMPAS FFSL path (the per-tracer FFSL work in
FP32 over the plain form's FP64 maps, streamlines, geometry and tracer
state) against the plain FP64 form (bit-exact with the oracle, see
test_gpu_us.py) on the same seeded meshes.  Tolerances: max |difference|
<= 1e-5 of each tracer's largest magnitude, relative mass drift <= 1e-6
against the FP64 form's mass."""
import sys

import numpy as np
import pytest

from ems_amd import mesh as MS, structs_us as U
from oracle import oracle_us as OU

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def errors(a, ref):
    rng = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-30)
    return (np.abs(a - ref) / rng).max(axis=1)


def pointwise(a, ref, floor=1e-3):
    big = np.abs(ref) >= floor * np.abs(ref).max(axis=1, keepdims=True)
    rel = np.where(big, np.abs(a - ref) / np.maximum(np.abs(ref), 1e-300), 0.0)
    return rel.max(axis=1)


def masses(case):
    return np.array([OU.tmass(case, n) for n in range(case.ntr)])


def compare(case, checkpoints, tol=1e-5):
    """Step an FP64 plain and an FP32 context side by side; the FFSL
    toggle (ffsl_init every other call) is the window's, so each context
    starts its steps from the same toggle state."""
    L = MS.lib()
    p = MS.Context(case, MS.PLAIN_F64)
    p.upload()
    q = MS.Context(case, MS.BATCH_F32)
    q.upload()
    done, out = 0, []
    for upto in checkpoints:
        f0 = case.wc.flag
        p.step(upto - done)
        p.sync()
        f1 = case.wc.flag
        case.wc.flag = f0
        q.step(upto - done)
        q.sync()
        assert case.wc.flag == f1
        done = upto
        p.download()
        ref = case.tracers()
        mp = masses(case)
        q.download()
        got = case.tracers()
        mq = masses(case)
        e, pw = errors(got, ref), pointwise(got, ref)
        drift = np.abs(mq - mp) / np.abs(mp)
        say("step %d: error %.3g (tracer %d), pointwise above 1e-3 %.3g, mass against FP64 "
            "%.3g" % (upto, e.max(), int(np.argmax(e)), pw.max(), drift.max()))
        assert np.isfinite(got).all()
        assert e.max() <= tol, (upto, e)
        assert drift.max() <= 1e-6, (upto, drift)
        out.append((e, pw, drift))
    p.close()
    q.close()
    L.emstr_us_forget_window(case.gp)
    return out


def test_hex_ffsl_fp32_vs_plain():
    """A 24 x 20 hexagon basin, 8 layers, 3 tracers, FFSL: 40 steps."""
    say("case hex ffsl fp32")
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=17, trasc=U.FFSL)
    case.wc.trsf = 0.8
    case.prepare(OU.set_dz)
    t0 = case.tracers()
    compare(case, (10, 40))
    assert np.abs(case.tracers() - t0).max() > 1e-3


def test_hex_ffsl_fp32_clip_vs_plain():
    """The FP32 form with FILL_METHOD CLIP against the plain form with it:
    40 steps."""
    say("case hex ffsl fp32 clip")
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=17, trasc=U.FFSL, dt=600.0)
    case.wc.trsf = 0.8
    case.wc.fillf = U.CLIP
    case.prepare(OU.set_dz)
    compare(case, (10, 40))


def test_hex_ffsl_fp32_multi_segment_vs_plain():
    """Streamlines of several segments (dt = 6000 s, up to 2 m/s: many
    cross two to four cells) in the FP32 form against the plain form: 20
    steps."""
    say("case hex ffsl fp32 multi-segment")
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=17, trasc=U.FFSL, dt=6000.0, vmax=2.0)
    case.prepare(OU.set_dz)
    compare(case, (5, 20))


def test_fp32_form_refuses_other_schemes():
    case = MS.hex_basin(nx=10, ny=8, nz=4, ntr=1, seed=3, trasc=U.VANLEER)
    L = MS.lib()
    assert not L.emstr_us_create_form(case.gp, case.wp, case.pp, MS.BATCH_F32)
    assert "FP32 form" in MS.last_error()[0]
    case = MS.hex_basin(nx=10, ny=8, nz=4, ntr=1, seed=3, trasc=U.FFSL)
    case.wc.rkstage = 2
    assert not L.emstr_us_create_form(case.gp, case.wp, case.pp, MS.BATCH_F32)
    assert "one Runge-Kutta stage" in MS.last_error()[0]
    assert not L.emstr_us_create_form(case.gp, case.wp, case.pp, 7)
    case.wc.rkstage = 1


@pytest.mark.parametrize("seed", [5, 3])
def test_c4_ffsl_fp32_vs_plain_500_steps(seed):
    """The C4 mesh (350 x 350 hexagons, 50 layers, 60 tracers, dt = 180 s)
    with FFSL: 500 steps of the FP32 form against the FP64 plain form."""
    say("case c4 ffsl fp32, 500 steps, seed", seed)
    case = MS.c4(seed=seed, trasc=U.FFSL)
    L = MS.lib()
    L.set_dz(case.gp, case.wp, case.pp)
    L.emstr_us_forget_window(case.gp)
    case.w_from_continuity()
    compare(case, (100, 500))
