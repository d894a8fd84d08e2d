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


# ---- set_dz_at_u1 / u2 and the offline transport step -------------------

def _eta_pattern(case, amp, phase=0.0):
    """A smooth elevation of amplitude amp (m) on the wet columns, with the
    lateral ghosts holding their interior value."""
    g = case.g
    w2 = S.view(g.w2_t, case.sgsizS, np.int32)[1:g.v2_t + 1]
    i = case.s2i[w2].astype(float)
    j = case.s2j[w2].astype(float)
    eta = case.a2(case.wd.eta)
    eta[w2] = amp * np.sin(0.35 * i + phase) * np.cos(0.27 * j - phase)
    return w2


def test_set_dz_at_u_column_geometry():
    """Per face column: the surface face is the layer holding the higher
    elevation either side (not below the face bed), the thicknesses from
    it to the bottom face telescope to (that elevation - face bed), the
    faces above carry the surface face's thickness, those below zero."""
    case = _case(seed=7)
    g = case.g
    dl = case.layers[-1] - case.layers[-2]
    _eta_pattern(case, 1.4 * dl)
    eta = case.a2(case.wd.eta)
    bpt = S.view(g.bpt, g.nbpt + 1, np.int32)[1:]
    bin_ = S.view(g.bin, g.nbpt + 1, np.int32)[1:]
    sm = bpt <= g.sgnumS
    eta[bpt[sm]] = eta[bin_[sm]]
    zm1 = S.view(g.zm1, case.sgsiz, np.int32)
    zp1 = S.view(g.zp1, case.sgsiz, np.int32)
    crossed = 0
    for d in (1, 2):
        O.set_dz_at_u(case, d)
        dzu = case.a3(case.wd.dzu1 if d == 1 else case.wd.dzu2)
        nb2 = g.b2_e1 if d == 1 else g.b2_e2
        w2 = S.view(g.w2_e1 if d == 1 else g.w2_e2, case.sgsizS, np.int32)[1:nb2 + 1]
        sur = S.view(g.sur_e1 if d == 1 else g.sur_e2, case.sgsizS, np.int32)[1:nb2 + 1]
        bot = S.view(g.bot_e1 if d == 1 else g.bot_e2, case.sgsizS, np.int32)[1:nb2 + 1]
        bz = case.a2(g.botzu1 if d == 1 else g.botzu2)
        nbm = S.view(g.xm1 if d == 1 else g.ym1, case.sgsiz, np.int32)
        for cs, su, cb in zip(w2, sur, bot):
            top = max(eta[cs], eta[nbm[cs]], bz[cs])
            k = case.s2k[su]
            assert case.layers[k] <= top and (k == case.nz - 1 or top < case.layers[k + 1])
            crossed += int(k < case.nz - 1)
            col = []
            c = su
            while True:
                col.append(c)
                if c == cb:
                    break
                c = zm1[c]
            assert abs(dzu[col].sum() - (top - bz[cs])) <= 1e-12 * max(1.0, top - bz[cs])
            c = su
            while c != cs:
                c = zp1[c]
                assert dzu[c] == dzu[su]
            c = zm1[cb]
            while zm1[c] != c:
                assert dzu[c] == 0.0
                c = zm1[c]
    assert crossed > 10


def _stored_flow(case, n, u0, v0, amp):
    """The stored flow of transport step n: velocities modulated in time,
    an elevation pattern moving in phase."""
    s = 1.0 + 0.3 * np.sin(0.9 * n)
    case.a3(case.wd.u1)[:] = u0 * s
    case.a3(case.wd.u2)[:] = v0 * (2.0 - s)
    _eta_pattern(case, amp, phase=0.4 * n)


def test_transport_step_steady_flow_is_the_online_step():
    """With the stored flow equal to the window's prepared fields (flat
    surface), an offline step recomputes the face thicknesses, fluxes and
    w that the prepared fields hold, so it equals the online tracer step
    to rounding."""
    case = _case(seed=9)
    case.wd.dttr = case.wd.dt
    t0 = case.tracers().copy()
    O.step(case)
    online = case.tracers().copy()
    case.set_tracers(t0)
    assert O.transport_step(case) == 0
    off = case.tracers()
    assert np.abs(online - t0).max() > 1e-3
    assert np.abs(off - online).max() <= 1e-12 * np.abs(online).max()


def test_transport_steps_conserve_mass_with_time_varying_flow():
    case = _case(seed=10)
    case.wd.dttr = case.wd.dt
    u0 = case.a3(case.wd.u1).copy()
    v0 = case.a3(case.wd.u2).copy()
    m0 = np.array([O.tmass(case, n) for n in range(case.ntr)])
    t0 = case.tracers().copy()
    for n in range(5):
        _stored_flow(case, n, u0, v0, 0.0)
        assert O.transport_step(case) == 0
    m1 = np.array([O.tmass(case, n) for n in range(case.ntr)])
    assert np.abs(case.tracers() - t0).max() > 1e-3
    assert np.all(np.abs(m1 - m0) <= 1e-12 * np.abs(m0))
