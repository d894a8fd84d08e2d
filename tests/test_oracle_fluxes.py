"""Pinning of the oracle's restatements of the flow-side inputs of the
tracer step: set_flux_3d (model/hd/momentum/vel3d.c:4041-4191) and the
continuity loop of vel_w_update (vel3d.c:4680-4711).

What fixes each, independently of the formulas restated there:
  set_flux_3d
      the seeded windows build their face velocities as volume flux /
      (face thickness x face length), so u dzu h x mdx must give the
      fluxes back to rounding; faces one cell deep take the 2D flux over
      the step divided by dt exactly, faces above a column's surface face
      are zero, and under FFSL the mean fluxes are copied
  vel_w_update
      the fluxes come from a corner streamfunction, so every column's
      depth-integrated transport is divergence-free: with w = 0 at the
      bed the upward flux through the bottom of each surface cell must
      cancel that cell's net lateral inflow; and w at every face equals
      the bed value plus the partial sums of the net inflows below it,
      summed here over the dense arrays in another order"""
import numpy as np

from ems_amd import configs, structs as S
from oracle import oracle as O


def _case(seed=3, trasc=S.VANLEER):
    case = configs.small(seed=seed, ntr=1, trasc=trasc)
    case.prepare(O.set_dz)
    return case


def _wet_faces(case, d):
    g = case.g
    w3 = S.view(g.w3_e1 if d == 1 else g.w3_e2, case.sgsiz, np.int32)
    nb = g.b3_e1 if d == 1 else g.b3_e2
    return w3[1:nb + 1]


def test_set_flux_3d_gives_the_face_fluxes_back():
    case = _case()
    f1 = case.a3(case.wd.u1flux3d).copy()
    f2 = case.a3(case.wd.u2flux3d).copy()
    case.a3(case.wd.u1flux3d)[:] = 7.0
    case.a3(case.wd.u2flux3d)[:] = 7.0
    O.set_flux_3d(case)
    for d, ref in ((1, f1), (2, f2)):
        got = case.a3(case.wd.u1flux3d if d == 1 else case.wd.u2flux3d)
        wf = _wet_faces(case, d)
        scale = np.abs(ref[wf]).max()
        assert scale > 0
        assert np.abs(got[wf] - ref[wf]).max() <= 1e-13 * scale
        other = np.ones(case.sgsiz, bool)
        other[wf] = False
        other[0] = False
        assert np.all(got[other] == 0.0)


def test_set_flux_3d_one_cell_deep_and_above_the_surface():
    case = _case(seed=4)
    g, wd, wc = case.g, case.wd, case.wc
    rng = np.random.default_rng(1)
    O.set_flux_3d(case)
    base = case.a3(wd.u1flux3d).copy()
    w2 = S.view(g.w2_e1, case.sgsizS, np.int32)[1:g.b2_e1 + 1]
    sur = S.view(g.sur_e1, case.sgsizS, np.int32)
    zm1 = S.view(g.zm1, case.sgsiz, np.int32)
    # lower the surface face of some columns by one layer
    pick = rng.choice(len(w2), 25, replace=False)
    lowered = []
    for p in pick:
        top = w2[p]
        below = zm1[top]
        if below != top and zm1[below] != below:
            sur[p + 1] = below
            lowered.append(top)
    # faces one cell deep with their 2D flux
    cb = S.view(wc.cbot_e1, case.sgsizS, np.int32)
    cand = [w2[p] for p in range(len(w2)) if p not in pick][:12]
    for i, f in enumerate(cand):
        cb[i + 1] = f
    wc.ncbot_e1 = len(cand)
    uf = case.a2(wd.u1flux)
    m2d = S.view(g.m2d, case.sgsiz, np.int32)
    uf[m2d[cand]] = rng.uniform(-50.0, 50.0, len(cand)) * case.dt
    O.set_flux_3d(case)
    got = case.a3(wd.u1flux3d)
    assert len(lowered) > 5
    assert np.all(got[lowered] == 0.0)
    assert np.array_equal(got[cand], uf[m2d[cand]] / case.dt)
    rest = np.ones(case.sgsiz, bool)
    rest[lowered] = rest[cand] = rest[0] = False
    assert np.array_equal(got[rest], base[rest])
    assert np.any(base[lowered] != 0.0)


def test_set_flux_3d_ffsl_copies_the_mean_fluxes():
    case = _case(seed=5)
    g, wd = case.g, case.wd
    rng = np.random.default_rng(2)
    vm = case.a3(wd.u1vm)
    vm[:] = rng.normal(0.0, 10.0, case.sgsiz)
    u1 = case.a3(wd.u1)
    u1[:] = 1.0
    case.wc.trasc = S.FFSL
    O.set_flux_3d(case)
    wf = _wet_faces(case, 1)
    got = case.a3(wd.u1flux3d)
    assert np.array_equal(got[wf], vm[wf])
    past = S.view(g.w3_e1, case.sgsiz, np.int32)[g.b3_e1 + 1:g.n3_e1 + 1]
    assert len(past) > 0
    assert np.all(u1[past] == 0.0) and np.all(vm[past] == 0.0)


def test_vel_w_update_closes_each_column():
    case = _case(seed=6)
    g, wd, wc = case.g, case.wd, case.wc
    O.set_flux_3d(case)
    O.vel_w_update(case)
    w = case.a3(wd.w)
    f1 = case.a3(wd.u1flux3d)
    f2 = case.a3(wd.u2flux3d)
    xp1 = S.view(g.xp1, case.sgsiz, np.int32)
    yp1 = S.view(g.yp1, case.sgsiz, np.int32)
    m2d = S.view(g.m2d, case.sgsiz, np.int32)
    area = case.a2(g.cellarea)
    s1 = S.view(wc.s1, case.sgsiz, np.int32)
    i1 = S.view(wc.i1, case.sgsizS, np.int32)
    wmax = np.abs(w).max()
    assert wmax > 0
    zp1 = S.view(g.zp1, case.sgsiz, np.int32)
    ncol = 0
    for cc in range(1, wc.vcs + 1):
        top, bed = s1[cc], i1[cc]
        if top == bed:
            continue
        # inflows of the cells from the bed up, summed in numpy order
        cells = [bed]
        while cells[-1] != top:
            cells.append(zp1[cells[-1]])
        cells = np.array(cells)
        inflow = f1[cells] - f1[xp1[cells]] + f2[cells] - f2[yp1[cells]]
        a = area[m2d[top]]
        part = np.concatenate([[0.0], np.cumsum(inflow[:-1])]) / a
        np.testing.assert_allclose(w[cells], part, rtol=0, atol=1e-12 * wmax)
        # divergence-free column: the surface cell's bottom flux cancels
        # its own net inflow
        assert abs(w[top] * a + inflow[-1]) <= 1e-10 * wmax * a
        ncol += 1
    assert ncol > 50
