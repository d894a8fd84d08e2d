"""GPU parity of the COMPAS plain FP64 form: the drop-in entry points of
libemstr_us.so (tracer_step_3d, advect_diffuse, vert_diffuse_3d, set_dz)
against the COMPAS CPU oracle on the same seeded meshes.  Tracer values,
cell lists and layer thicknesses must agree exactly."""
import sys

import numpy as np
import pytest

from ems_amd import configs, mesh as MS, structs as S, structs_us as U
from oracle import oracle as O, oracle_us as OU
from tests import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_both(case, nsteps, which="step"):
    L = MS.lib()
    snap = H.snapshot(case)
    for _ in range(nsteps):
        if which == "step":
            L.tracer_step_3d(case.gp, case.wp, case.pp)
        elif which == "adv":
            assert L.advect_diffuse(case.gp, case.wp, case.pp) == 0, MS.last_error()
        else:
            L.vert_diffuse_3d(case.gp, case.wp, case.pp)
        msg, cell = MS.last_error()
        assert msg == "", (msg, cell)
    gpu = case.tracers()
    b1_gpu = case.wc.b1
    H.restore(case, snap)
    for _ in range(nsteps):
        if which == "step":
            assert OU.step(case) == 0
        elif which == "adv":
            assert OU.advect_diffuse(case) == 0
        else:
            assert OU.vert_diffuse(case) == 0
    ora = case.tracers()
    assert b1_gpu == case.wc.b1
    L.emstr_us_forget_window(case.gp)
    return gpu, ora


def assert_exact(gpu, ora):
    d = np.abs(gpu - ora)
    scale = np.maximum(np.abs(ora), 1e-300)
    rel = (d / scale).max()
    assert np.array_equal(gpu, ora) or rel <= 1e-12, ("max rel diff", rel, int((d > 0).sum()))


CASES = [
    ("vanleer", dict(trasc=U.VANLEER), U.NONE, 4),
    ("vanleer+ultimate", dict(trasc=U.VANLEER, ultimate=1), U.NONE, 4),
    ("order1", dict(trasc=U.ORDER1), U.NONE, 4),
    ("quickest", dict(trasc=U.QUICKEST), U.NONE, 4),
    ("quickest+ultimate", dict(trasc=U.QUICKEST, ultimate=1), U.NONE, 4),
    ("vanleer_substeps", dict(trasc=U.VANLEER, dt=3600.0, vmax=1.2), U.SUB_STEP, 2),
]


@pytest.mark.parametrize("name,kw,stab,nsteps", CASES)
def test_hex_step_bit_exact(name, kw, stab, nsteps):
    say("case hex", name)
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=7, **kw)
    case.wc.stab = stab
    case.prepare(OU.set_dz)
    t0 = case.tracers()
    gpu, ora = run_both(case, nsteps)
    assert np.abs(ora - t0).max() > 1e-3
    if stab == U.SUB_STEP:
        assert case.wc.b1 < case.dt
    assert_exact(gpu, ora)


@pytest.mark.parametrize("trasc", [S.VANLEER, S.QUICKEST])
def test_quad_mesh_step_bit_exact(trasc):
    say("case quads of a structured window", trasc)
    sc = configs.small(seed=21, ntr=3, dt=120.0, trasc=trasc, ultimate=1)
    sc.prepare(O.set_dz)
    qc = MS.quads_from_case(sc)
    OU.set_dz(qc.gp, qc.wp, qc.pp)
    gpu, ora = run_both(qc, 3)
    assert_exact(gpu, ora)


def test_entry_points_and_diffusion():
    say("case hex adv / vdiff entry points, horizontal diffusion")
    case = MS.hex_basin(nx=20, ny=16, nz=7, ntr=2, seed=11)
    U.view(case.wc.diffuse, 2, np.int32)[:] = [1, 0]
    rng = np.random.default_rng(0)
    case.ae(case.wc.u2kh)[:] = rng.uniform(0.0, 50.0, case.sze)
    case.prepare(OU.set_dz)
    gpu, ora = run_both(case, 2, "adv")
    assert_exact(gpu, ora)
    gpu, ora = run_both(case, 2, "vdiff")
    assert_exact(gpu, ora)
    gpu, ora = run_both(case, 2, "step")
    assert_exact(gpu, ora)


def _moving(case, seed):
    rng = np.random.default_rng(seed)
    g = case.g
    w2 = U.view(g.w2_t, case.szcS, np.int32)[1:g.v2_t + 1]
    dl = case.layers[1] - case.layers[0]
    old = case.a2(case.wc.oldeta)
    eta = case.a2(case.wd.eta)
    old[w2] = rng.uniform(-1.3, 0.0, len(w2)) * dl
    eta[w2] = old[w2]
    return w2, old, eta, rng, dl


def test_moving_surface_step():
    say("case hex moving surface")
    case = MS.hex_basin(nx=20, ny=16, nz=6, ntr=3, seed=13, dmin=2.0, dmax=30.0)
    w2, old, eta, rng, dl = _moving(case, 4)
    OU.set_dz(case.gp, case.wp, case.pp)
    eta[w2] = old[w2] + rng.uniform(-0.9, 0.9, len(w2)) * dl
    case.prepare(OU.set_dz)
    gpu, ora = run_both(case, 1)
    assert_exact(gpu, ora)


def test_set_dz_lists_bit_exact():
    say("case hex set_dz lists")
    mk = lambda: MS.hex_basin(nx=16, ny=14, nz=7, ntr=1, seed=31, dmin=2.0, dmax=40.0)  # noqa: E731
    a, b = mk(), mk()
    wa = _moving(a, 7)
    wb = _moving(b, 7)
    L = MS.lib()
    for step in range(2):
        for c, (w2, old, eta, rng, dl) in ((a, wa), (b, wb)):
            if step:
                eta[w2] = old[w2] + np.random.default_rng(9).uniform(-1.0, 1.0, len(w2)) * dl
        L.set_dz(a.gp, a.wp, a.pp)
        assert MS.last_error()[0] == ""
        OU.set_dz(b.gp, b.wp, b.pp)
        for f in ("vc", "vcs", "vc1", "vcs1"):
            assert getattr(a.wc, f) == getattr(b.wc, f), f
        vc, vcs = b.wc.vc, b.wc.vcs
        assert np.array_equal(U.view(a.wc.s1, a.szc, np.int32)[1:vc + 1],
                              U.view(b.wc.s1, b.szc, np.int32)[1:vc + 1])
        for f in ("i1", "i2", "i3"):
            assert np.array_equal(U.view(getattr(a.wc, f), a.szcS, np.int32)[1:vcs + 1],
                                  U.view(getattr(b.wc, f), b.szcS, np.int32)[1:vcs + 1]), f
        n2 = a.g.v2_t
        for f in ("sur_t", "nsur_t"):
            assert np.array_equal(U.view(getattr(a.g, f), a.szcS, np.int32)[1:n2 + 1],
                                  U.view(getattr(b.g, f), b.szcS, np.int32)[1:n2 + 1]), f
        assert np.array_equal(a.ac(a.wc.dz)[a.wet], b.ac(b.wc.dz)[b.wet])
    L.emstr_us_forget_window(a.gp)


def test_unsupported_settings_are_reported():
    say("case hex unsupported")
    case = MS.hex_basin(nx=10, ny=8, nz=4, ntr=1, seed=2)
    case.prepare(OU.set_dz)
    case.wc.rkstage = 2
    MS.lib().tracer_step_3d(case.gp, case.wp, case.pp)
    assert "Runge-Kutta" in MS.last_error()[0]
    case.wc.rkstage = 1
    MS.lib().emstr_us_forget_window(case.gp)
