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
    """The GPU run, then the oracle from the same state: tracers, and for
    FFSL the ffsl_init toggle and the centred velocities it writes."""
    L = MS.lib()
    snap = H.snapshot(case)
    flag0 = case.wc.flag
    uv0 = case.ac(case.wd.u).copy(), case.ac(case.wd.v).copy()
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
    flag_gpu, uv_gpu = case.wc.flag, (case.ac(case.wd.u).copy(), case.ac(case.wd.v).copy())
    H.restore(case, snap)
    case.wc.flag = flag0
    case.ac(case.wd.u)[:], case.ac(case.wd.v)[:] = uv0
    for _ in range(nsteps):
        if which == "step":
            assert OU.step(case) == 0
        elif which == "adv":
            assert OU.advect_diffuse(case) == 0
        else:
            assert OU.vert_diffuse(case) == 0
    ora = case.tracers()
    assert b1_gpu == case.wc.b1
    assert flag_gpu == case.wc.flag
    assert np.array_equal(uv_gpu[0], case.ac(case.wd.u))
    assert np.array_equal(uv_gpu[1], case.ac(case.wd.v))
    L.emstr_us_forget_window(case.gp)
    return gpu, ora


def assert_exact(gpu, ora):
    d = np.abs(gpu - ora)
    scale = np.maximum(np.abs(ora), 1e-300)
    rel = (d / scale).max()
    assert np.array_equal(gpu, ora), ("max rel diff", rel, int((d > 0).sum()))


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


@pytest.mark.parametrize("trasc,ult", [(U.VANLEER, 0), (U.VANLEER, 1), (U.QUICKEST, 1)])
def test_zero_transport_and_flat_fields(trasc, ult):
    """Edges and faces without transport, and tracer plateaus where the
    limiter's product is exactly zero or neighbours are equal: the faces
    (branch-free van Leer with the mirrored stencil) keep the reference's
    values."""
    say("case hex zero transport, plateaus", trasc, ult)
    case = MS.hex_basin(nx=20, ny=16, nz=6, ntr=3, seed=29, trasc=trasc, ultimate=ult)
    rng = np.random.default_rng(5)
    u1, f = case.ae(case.wd.u1), case.ae(case.wd.u1flux3d)
    still = rng.random(u1.shape) < 0.3
    u1[still] = 0.0
    f[still] = 0.0
    w = case.ac(case.wd.w)
    w[rng.random(w.shape) < 0.3] = 0.0
    case.prepare(OU.set_dz)
    tr = case.tracers()
    tr[0] = np.round(tr[0])
    tr[1] = np.where(rng.random(tr[1].shape) < 0.5, 3.0, tr[1])
    case.set_tracers(tr)
    gpu, ora = run_both(case, 3)
    assert np.abs(ora - tr).max() > 1e-3
    assert np.array_equal(gpu, ora), np.abs(gpu - ora).max()


RK_CASES = [
    # name, hex_basin kw, rkstage, stab, nsteps
    ("vanleer_rk2", dict(trasc=U.VANLEER), 2, U.NONE, 3),
    ("vanleer_rk3", dict(trasc=U.VANLEER), 3, U.NONE, 3),
    ("quickest_ultimate_rk3", dict(trasc=U.QUICKEST, ultimate=1), 3, U.NONE, 3),
    ("order1_rk2", dict(trasc=U.ORDER1), 2, U.NONE, 3),
    ("vanleer_rk3_substeps", dict(trasc=U.VANLEER, dt=3600.0, vmax=1.2), 3, U.SUB_STEP, 2),
    ("ffsl_rk2", dict(trasc=U.FFSL), 2, U.NONE, 3),
    ("ffsl_rk3", dict(trasc=U.FFSL), 3, U.NONE, 2),
    ("vanleer_rk5", dict(trasc=U.VANLEER), 5, U.NONE, 2),
    ("quickest_ultimate_rk5", dict(trasc=U.QUICKEST, ultimate=1), 5, U.NONE, 2),
    ("vanleer_rk5_substeps", dict(trasc=U.VANLEER, dt=3600.0, vmax=1.2), 5, U.SUB_STEP, 1),
    ("ffsl_rk5", dict(trasc=U.FFSL), 5, U.NONE, 2),
]


@pytest.mark.parametrize("name,kw,rkstage,stab,nsteps", RK_CASES)
def test_hex_runge_kutta_stages(name, kw, rkstage, stab, nsteps):
    """Runge-Kutta stage loop (tracers.c:841-1098): 2, 3 and 5 stages
    against the oracle (whose stages are pinned by the forward-Euler
    compositions of tests/test_oracle_us_rk.py), every scheme bit for bit."""
    say("case hex", name)
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=7, **kw)
    case.wc.rkstage = rkstage
    case.wc.stab = stab
    case.wc.trsf = 0.8
    case.prepare(OU.set_dz)
    t0 = case.tracers()
    gpu, ora = run_both(case, nsteps)
    assert np.abs(ora - t0).max() > 1e-3
    if stab == U.SUB_STEP:
        assert case.wc.b1 < case.dt
    assert_exact(gpu, ora)
    case.wc.rkstage = 1


@pytest.mark.parametrize("rkstage", [2, 3, 5])
def test_moving_surface_runge_kutta(rkstage):
    """A moving surface with stages: the surface cell's mass at the start of
    the sub-step is recomputed from its buffer at every stage."""
    say("case hex moving surface, rkstage", rkstage)
    case = MS.hex_basin(nx=20, ny=16, nz=6, ntr=3, seed=13, dmin=2.0, dmax=30.0)
    w2, old, eta, rng, dl = _moving(case, 4)
    OU.set_dz(case.gp, case.wp, case.pp)
    eta[w2] = old[w2] + rng.uniform(-0.9, 0.9, len(w2)) * dl
    case.wc.rkstage = rkstage
    case.prepare(OU.set_dz)
    gpu, ora = run_both(case, 1)
    assert_exact(gpu, ora)
    case.wc.rkstage = 1


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


FFSL_CASES = [
    # name, hex_basin kw, stab, trsf, nsteps
    ("ffsl", dict(), U.NONE, 0.8, 4),
    ("ffsl_substeps", dict(dt=120.0), U.SUB_STEP, 0.002, 3),
    ("ffsl_fast", dict(dt=300.0, vmax=1.2), U.NONE, 0.8, 3),
]


@pytest.mark.parametrize("name,kw,stab,trsf,nsteps", FFSL_CASES)
def test_hex_ffsl_step(name, kw, stab, trsf, nsteps):
    """FFSL streamlines, least squares and transverse terms; ffsl_init
    runs on every other call, so several steps cover both branches.  The
    streamline directions are u / |u| and v / |u| on both sides (no math
    library), so tracers agree bit for bit."""
    say("case hex", name)
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=17, trasc=U.FFSL, **kw)
    case.wc.stab = stab
    case.wc.trsf = trsf
    case.prepare(OU.set_dz)
    t0 = case.tracers()
    flag0 = case.wc.flag
    gpu, ora = run_both(case, nsteps)
    assert np.abs(ora - t0).max() > 1e-3
    if stab == U.SUB_STEP:
        assert case.wc.b1 < case.dt
    assert case.wc.flag == (flag0 ^ (U.F_FFSL if nsteps % 2 else 0))
    assert_exact(gpu, ora)


CLIP_CASES = [
    # name, hex_basin kw, stab, trsf, nsteps, fill
    ("clip", dict(dt=600.0), U.NONE, 0.8, 3, U.CLIP),
    ("clip_monotonic", dict(dt=600.0), U.NONE, 0.8, 2, U.CLIP | U.MONOTONIC),
    ("clip_substeps", dict(dt=120.0), U.SUB_STEP, 0.002, 3, U.CLIP),
]


@pytest.mark.parametrize("name,kw,stab,trsf,nsteps,fill", CLIP_CASES)
def test_hex_ffsl_clip(name, kw, stab, trsf, nsteps, fill):
    """FILL_METHOD CLIP (clip_ffsl): each centre clamped to the range of
    its streamlines' stencils; tracers bit for bit with the oracle, and
    the clip acts (the result differs from the unclipped step)."""
    say("case hex", name)
    res = {}
    for f in (0, fill):
        case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=17, trasc=U.FFSL, **kw)
        case.wc.stab = stab
        case.wc.trsf = trsf
        case.wc.fillf = f
        case.prepare(OU.set_dz)
        gpu, ora = run_both(case, nsteps)
        assert_exact(gpu, ora)
        res[f] = gpu
    assert (res[0] != res[fill]).sum() > 50


def test_hex_ffsl_clip_moving_surface():
    """The surface moves across layers in some columns, and every column
    stays processed (the clip is refused otherwise)."""
    say("case hex ffsl clip, moving surface + diffusion")
    case = MS.hex_basin(nx=20, ny=16, nz=6, ntr=3, seed=19, dmin=6.0, dmax=30.0,
                        trasc=U.FFSL, dt=300.0)
    rng = np.random.default_rng(5)
    g = case.g
    w2 = U.view(g.w2_t, case.szcS, np.int32)[1:g.v2_t + 1]
    dl = case.layers[1] - case.layers[0]
    old, eta = case.a2(case.wc.oldeta), case.a2(case.wd.eta)
    old[w2] = rng.uniform(-0.7, 0.0, len(w2)) * dl
    eta[w2] = old[w2] + rng.uniform(-0.5, 0.5, len(w2)) * dl
    U.view(case.wc.diffuse, 3, np.int32)[:] = [1, 0, 1]
    case.ae(case.wc.u2kh)[:] = rng.uniform(0.0, 30.0, case.sze)
    case.wc.fillf = U.CLIP
    case.prepare(OU.set_dz)
    gpu, ora = run_both(case, 2)
    assert_exact(gpu, ora)


def test_hex_ffsl_clip_settings_outside_scope():
    """MONOTONIC without CLIP (clamps to bounds the reference never sets),
    CLIP with two Runge-Kutta stages, CLIP with a tracer not advected."""
    L = MS.lib()
    for fill, rk, adv, what in ((U.MONOTONIC, 1, 1, "MONOTONIC"), (U.CLIP, 2, 1, "one Runge"),
                                (U.CLIP, 1, 0, "advected")):
        say("case hex ffsl clip refused:", what)
        case = MS.hex_basin(nx=12, ny=10, nz=4, ntr=2, seed=2, trasc=U.FFSL, dt=300.0)
        case.wc.fillf = fill
        case.wc.rkstage = rk
        U.view(case.wc.advect, 2, np.int32)[1] = adv
        case.prepare(OU.set_dz)
        L.tracer_step_3d(case.gp, case.wp, case.pp)
        assert what in MS.last_error()[0], MS.last_error()
        L.emstr_us_forget_window(case.gp)
    say("case hex ffsl clip refused: a column not processed")
    case = MS.hex_basin(nx=20, ny=16, nz=6, ntr=3, seed=19, dmin=2.0, dmax=30.0,
                        trasc=U.FFSL, dt=300.0)
    w2, old, eta, rng, dl = _moving(case, 5)
    OU.set_dz(case.gp, case.wp, case.pp)
    eta[w2] = old[w2] + rng.uniform(-0.9, 0.9, len(w2)) * dl
    case.wc.fillf = U.CLIP
    case.prepare(OU.set_dz)
    assert case.wc.vcs1 < case.g.v2_t
    L.tracer_step_3d(case.gp, case.wp, case.pp)
    assert "not processed" in MS.last_error()[0], MS.last_error()
    L.emstr_us_forget_window(case.gp)


def test_hex_ffsl_vertical_courant_above_one():
    """A vertical velocity 300 times that of continuity: vertical Courant
    numbers up to 3.5, so the vertical fluxes read tr_moc (the mean
    streamline end value) in the fractional part and over whole layers,
    and the GPU computes tr_moc only where it is read."""
    say("case hex ffsl, w x 300")
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=17, trasc=U.FFSL, dt=300.0)
    case.wc.trsf = 0.8
    case.prepare(OU.set_dz)
    w = case.ac(case.wd.w)
    w *= 300.0
    wet = case.wet
    cr = np.abs(w[wet]) * case.dt / np.maximum(case.ac(case.wc.dz)[wet], 1e-9)
    assert (cr > 1.0).sum() >= 10
    gpu, ora = run_both(case, 2)
    assert np.isfinite(ora).all()
    assert_exact(gpu, ora)


def test_hex_ffsl_moving_surface_diffusion():
    say("case hex ffsl moving surface + diffusion")
    case = MS.hex_basin(nx=20, ny=16, nz=6, ntr=3, seed=19, dmin=2.0, dmax=30.0,
                        trasc=U.FFSL)
    w2, old, eta, rng, dl = _moving(case, 5)
    OU.set_dz(case.gp, case.wp, case.pp)
    eta[w2] = old[w2] + rng.uniform(-0.9, 0.9, len(w2)) * dl
    U.view(case.wc.diffuse, 3, np.int32)[:] = [1, 0, 1]
    case.ae(case.wc.u2kh)[:] = rng.uniform(0.0, 30.0, case.sze)
    case.prepare(OU.set_dz)
    gpu, ora = run_both(case, 2)
    assert_exact(gpu, ora)


def test_quad_mesh_ffsl_step():
    say("case quads ffsl")
    sc = configs.small(seed=23, ntr=2, dt=120.0, trasc=S.VANLEER)
    sc.prepare(O.set_dz)
    qc = MS.quads_from_case(sc)
    qc.set_tangential()
    qc.scheme(U.FFSL, 0, 120.0)
    OU.set_dz(qc.gp, qc.wp, qc.pp)
    gpu, ora = run_both(qc, 3)
    assert_exact(gpu, ora)


def _maps_gpu_vs_oracle(case):
    """One step on the GPU and in the oracle from the same state; the
    FFSL maps of the step (upstream centre of every wet edge, departure
    centres of faces and centres) must be identical."""
    L = MS.lib()
    snap = H.snapshot(case)
    flag0 = case.wc.flag
    uv0 = case.ac(case.wd.u).copy(), case.ac(case.wd.v).copy()
    L.tracer_step_3d(case.gp, case.wp, case.pp)
    assert MS.last_error()[0] == "", MS.last_error()
    gcl, gcf, gcc = MS.ffsl_maps(case)
    L.emstr_us_forget_window(case.gp)
    H.restore(case, snap)
    case.wc.flag = flag0
    case.ac(case.wd.u)[:], case.ac(case.wd.v)[:] = uv0
    assert OU.step(case) == 0
    ocl, ocf, occ = OU.ffsl_maps(case)
    g = case.g
    wet_e = U.view(g.w3_e1, case.sze, np.int32)[1:g.a3_e1 + 1]
    s1 = U.view(case.wc.s1, case.szc, np.int32)[1:case.wc.vc + 1]
    assert np.all(gcl[wet_e] > 0)
    assert np.array_equal(gcl[wet_e], ocl[wet_e])
    assert np.array_equal(gcf[s1], ocf[s1])
    assert np.array_equal(gcc[case.wet], occ[case.wet])
    # the maps are not trivial: some streamlines start in e2c[e][1]
    # (positive u1) and some departure centres are not the cell itself
    assert np.any(gcl[wet_e] == case.e2c[wet_e, 1]) and np.any(gcl[wet_e] == case.e2c[wet_e, 0])
    assert np.any(gcf[s1] != s1)


def test_ffsl_maps_bit_exact_hex():
    say("case hex ffsl maps")
    case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=1, seed=29, trasc=U.FFSL, dt=300.0)
    case.prepare(OU.set_dz)
    _maps_gpu_vs_oracle(case)


def test_ffsl_maps_bit_exact_c4():
    """The C4 mesh (350 x 350 hexagons, 50 layers), one tracer: the maps
    do not depend on the tracer."""
    say("case c4 ffsl maps (mesh build ~30 s)")
    case = MS.c4(seed=3, ntr=1, trasc=U.FFSL)
    case.prepare(OU.set_dz)
    say("c4 built:", case.nwet, "wet cells")
    _maps_gpu_vs_oracle(case)


@pytest.mark.parametrize("trasc", [U.FFSL, U.VANLEER])
def test_c4_tracer_values_vs_oracle(trasc):
    """The full C4 mesh (350 x 350 jittered hexagons, 50 layers), three
    tracers, two steps through the drop-in tracer_step_3d against the
    oracle: every tracer value bit for bit, FFSL and van Leer."""
    say("case c4 tracer values, trasc", trasc, "(mesh build ~30 s)")
    case = MS.c4(seed=3, ntr=3, trasc=trasc)
    case.prepare(OU.set_dz)
    say("c4 built:", case.nwet, "wet cells")
    start = case.tracers().copy()
    gpu, ora = run_both(case, 2)
    assert_exact(gpu, ora)
    assert not np.array_equal(gpu, start)


MULTI_CASES = [
    # name, hex_basin kw, FILL_METHOD, steps
    ("multi", dict(dt=6000.0, vmax=2.0), 0, 3),
    ("multi_long", dict(dt=12000.0, vmax=1.2), 0, 2),
    ("multi_clip", dict(dt=6000.0, vmax=2.0), U.CLIP, 3),
]


@pytest.mark.parametrize("name,kw,fill,nsteps", MULTI_CASES)
def test_hex_ffsl_multi_segment(name, kw, fill, nsteps):
    """Streamlines that cross into wet cells go on from the crossing with
    the velocity fitted there (hd_trans_interp), over steps where
    ffsl_init rebuilds the fit and steps where it keeps it: tracers bit
    for bit with the oracle; then one more step of each: the cell where
    every streamline ends is the same, and many end beyond the cells
    either side of their edge."""
    say("case hex ffsl", name)
    case = MS.hex_basin(nx=24, ny=20, nz=6, ntr=3, seed=17, trasc=U.FFSL, **kw)
    case.wc.fillf = fill
    case.prepare(OU.set_dz)
    t0 = case.tracers()
    gpu, ora = run_both(case, nsteps)
    assert_exact(gpu, ora)
    assert np.abs(ora - t0).max() > 1e-3
    L = MS.lib()
    snap = H.snapshot(case)
    # a new context: ffsl_init runs on its first call (it builds the fit)
    case.wc.flag &= ~U.F_FFSL
    flag0 = case.wc.flag
    uv0 = case.ac(case.wd.u).copy(), case.ac(case.wd.v).copy()
    L.tracer_step_3d(case.gp, case.wp, case.pp)
    assert MS.last_error()[0] == "", MS.last_error()
    gcl = MS.ffsl_maps(case)[0]
    L.emstr_us_forget_window(case.gp)
    H.restore(case, snap)
    case.wc.flag = flag0
    case.ac(case.wd.u)[:], case.ac(case.wd.v)[:] = uv0
    assert OU.step(case) == 0
    ocl = OU.ffsl_maps(case)[0]
    g = case.g
    wet_e = U.view(g.w3_e1, case.sze, np.int32)[1:g.a3_e1 + 1]
    assert np.array_equal(gcl[wet_e], ocl[wet_e])
    beyond = (gcl[wet_e] != case.e2c[wet_e, 0]) & (gcl[wet_e] != case.e2c[wet_e, 1])
    say("streamlines ending beyond their edge's cells:", int(beyond.sum()), "of", wet_e.size)
    assert beyond.sum() > 50


def test_ffsl_multi_segment_needs_a_velocity_fit():
    """A streamline that goes on into a wet cell before ffsl_init has run
    on the window (its toggle already set on the first call) has no
    velocity fit: reported."""
    say("case hex ffsl multi-segment, no fit")
    case = MS.hex_basin(nx=12, ny=10, nz=3, ntr=1, seed=2, trasc=U.FFSL, dt=12000.0)
    case.prepare(OU.set_dz)
    case.wc.flag |= U.F_FFSL
    MS.lib().tracer_step_3d(case.gp, case.wp, case.pp)
    assert "before ffsl_init" in MS.last_error()[0]
    MS.lib().emstr_us_forget_window(case.gp)


def test_unsupported_settings_are_reported():
    say("case hex unsupported")
    case = MS.hex_basin(nx=10, ny=8, nz=4, ntr=1, seed=2)
    case.prepare(OU.set_dz)
    case.wc.rkstage = 4
    MS.lib().tracer_step_3d(case.gp, case.wp, case.pp)
    assert "Runge-Kutta" in MS.last_error()[0]
    case.wc.rkstage = 1
    MS.lib().emstr_us_forget_window(case.gp)
