"""Pinning of the COMPAS oracle (oracle/oracle_us.c).

QUICKEST with ULTIMATE is where the limiter acts (van Leer is TVD, so
ULTIMATE barely changes it); the quadrilateral equivalence covers both.

The strongest check: on the quadrilateral mesh of a structured window the
COMPAS tracer step (edge stencils through c2c/e2e, edge signs eSc, the
Runge-Kutta update, surf_concrk) must equal the SHOC oracle's step, which
is itself pinned by closed forms (tests/test_oracle.py).  They differ
only in summation order, so the bar is 1e-12 relative.  On jittered
hexagons: constancy, closed-basin mass conservation, the set_dz column
walk and the mesh maps."""
import numpy as np
import pytest

from ems_amd import configs, mesh as MS, structs as S, structs_us as U
from oracle import oracle as O, oracle_us as OU


def _on_quads(sc, qc, ts):
    kk, jj, ii = qc.shoc_dense
    T = np.zeros((sc.ntr, sc.nz * sc.ny * sc.nx))
    T[:, sc.dense_index(sc.wet)] = ts
    return T.reshape(sc.ntr, sc.nz, sc.ny, sc.nx)[:, kk, jj, ii]


def _rel(a, b):
    return (np.abs(a - b) / np.maximum(np.abs(b), 1e-300)).max()


@pytest.mark.parametrize("trasc,ult", [(S.VANLEER, 0), (S.VANLEER, 1), (S.ORDER1, 0),
                                      (S.QUICKEST, 0), (S.QUICKEST, 1)])
def test_quad_mesh_step_equals_structured_step(trasc, ult):
    sc = configs.small(seed=21, ntr=3, dt=120.0, trasc=trasc, ultimate=ult)
    sc.prepare(O.set_dz)
    qc = MS.quads_from_case(sc)
    OU.set_dz(qc.gp, qc.wp, qc.pp)
    assert (sc.wc.vc, sc.wc.vcs) == (qc.wc.vc, qc.wc.vcs)
    t0 = qc.tracers()
    for _ in range(4):
        assert O.step(sc) == 0
        assert OU.step(qc) == 0
    got = qc.tracers()
    assert np.abs(got - t0).max() > 1e-3
    assert _rel(got, _on_quads(sc, qc, sc.tracers())) < 1e-12


def test_quad_mesh_moving_surface_and_diffusion():
    sc = configs.small(seed=13, ntr=2, nz=6, dmin=2.0, dmax=30.0, dt=120.0)
    rng = np.random.default_rng(4)
    g = sc.g
    w2 = S.view(g.w2_t, sc.sgsizS, np.int32)[1:g.v2_t + 1]
    dl = sc.layers[1] - sc.layers[0]
    old = sc.a2(sc.wc.oldeta)
    eta = sc.a2(sc.wd.eta)
    old[w2] = rng.uniform(-1.3, 0.0, len(w2)) * dl
    eta[w2] = old[w2]
    O.set_dz(sc.gp, sc.wp, sc.pp)
    eta[w2] = old[w2] + rng.uniform(-0.9, 0.9, len(w2)) * dl
    S.view(sc.wc.diffuse, 2, np.int32)[:] = [1, 0]
    kh = rng.uniform(0.0, 50.0, sc.sgsiz)
    sc.a3(sc.wc.u1kh)[:] = kh
    sc.a3(sc.wc.u2kh)[:] = kh[::-1]
    sc.prepare(O.set_dz)
    # the COMPAS window on the same state: old surface first, then new
    eta_new = eta.copy()
    eta[w2] = old[w2]
    qc = MS.quads_from_case(sc)
    eta[:] = eta_new
    q2 = qc.a2(qc.wc.oldeta)
    qe = qc.a2(qc.wd.eta)
    cs_q = U.view(qc.g.w2_t, qc.szcS, np.int32)[1:qc.g.v2_t + 1]
    m = qc.s2i[cs_q]
    # surface values by (i, j)
    top = {(sc.s2i[c], sc.s2j[c]): c for c in w2}
    src = np.array([top[(mm % sc.nx, mm // sc.nx)] for mm in m])
    q2[cs_q] = old[src]
    qe[cs_q] = old[src]
    OU.set_dz(qc.gp, qc.wp, qc.pp)
    qe[cs_q] = eta[src]
    OU.set_dz(qc.gp, qc.wp, qc.pp)
    # diffusivity at each COMPAS edge from the SHOC face it stands for
    U.view(qc.wc.diffuse, 2, np.int32)[:] = [1, 0]
    e = np.arange(1, qc.sze)
    mm, j = np.divmod(qc.e2m[e], 4)
    k = qc.s2k[qc.e2c[e, 0]]
    nx, ny, nz = sc.nx, sc.ny, sc.nz
    KH1 = np.zeros(nz * ny * nx)
    KH2 = np.zeros(nz * ny * nx)
    KH1[sc.dense_index(sc.wet)] = sc.a3(sc.wc.u1kh)[sc.wet]
    KH2[sc.dense_index(sc.wet)] = sc.a3(sc.wc.u2kh)[sc.wet]
    KH1, KH2 = KH1.reshape(nz, ny, nx), KH2.reshape(nz, ny, nx)
    i0, j0 = mm % nx, mm // nx
    kh_e = np.where(j == 0, KH1[k, j0, np.minimum(i0 + 1, nx - 1)],
                    np.where(j == 1, KH2[k, np.minimum(j0 + 1, ny - 1), i0],
                             np.where(j == 2, KH1[k, j0, i0], KH2[k, j0, i0])))
    qc.ae(qc.wc.u2kh)[e] = kh_e
    assert (sc.wc.vc, sc.wc.vcs) == (qc.wc.vc, qc.wc.vcs)
    assert O.step(sc) == 0
    assert OU.step(qc) == 0
    assert _rel(qc.tracers(), _on_quads(sc, qc, sc.tracers())) < 1e-12


@pytest.mark.parametrize("trasc,ult,stab", [(U.VANLEER, 0, U.NONE), (U.VANLEER, 1, U.NONE),
                                            (U.QUICKEST, 1, U.NONE),
                                            (U.ORDER1, 0, U.NONE),
                                            (U.VANLEER, 0, U.SUB_STEP)])
def test_hex_constancy_and_mass(trasc, ult, stab):
    case = MS.hex_basin(nx=20, ny=18, nz=7, ntr=2, trasc=trasc, ultimate=ult, seed=3,
                        dt=3600.0 if stab == U.SUB_STEP else 60.0, vmax=1.2)
    case.wc.stab = stab
    U.view(case.wc.mintr, 2, np.float64)[:] = -1e10
    c = case.tracers()
    c[0] = 3.25
    case.set_tracers(c)
    case.prepare(OU.set_dz)
    m0 = np.array([OU.tmass(case, n) for n in range(case.ntr)])
    for _ in range(5):
        assert OU.step(case) == 0
    if stab == U.SUB_STEP:
        assert case.wc.b1 < case.dt
    t = case.tracers()
    m1 = np.array([OU.tmass(case, n) for n in range(case.ntr)])
    assert np.abs(t[0] - 3.25).max() < 1e-11
    assert np.all(np.abs(m1 - m0) <= 1e-12 * np.abs(m0))


def test_hex_maps_are_consistent():
    case = MS.hex_basin(nx=12, ny=10, nz=5, ntr=1, seed=2)
    g = case.g
    npem = 6
    c2c = U.view2(g.c2c, npem + 1, case.szc)
    c2e = U.view2(g.c2e, npem + 1, case.szc)
    e2c = case.e2c
    e2e = U.view2(g.e2e, case.sze, 2)
    eSc = U.view2(g.eSc, npem + 1, case.szcS)
    m2d = U.view(g.m2d, case.szc, np.int32)
    wetset = set(case.wet.tolist())
    for c in case.wet:
        cs = m2d[c]
        for j in range(1, 7):
            e = c2e[j, c]
            assert e > 0
            assert c in (e2c[e, 0], e2c[e, 1])
            side = 0 if e2c[e, 0] == c else 1
            assert e2e[e, side] == j
            assert c2c[j, c] == e2c[e, 1 - side]
            # the 2D sign holds for edges between wet centres (a land
            # edge at depth is oriented from its wet side and carries no
            # flux)
            if c2c[j, c] in wetset:
                assert eSc[j, cs] == (1 if side == 1 else -1)
            # the ghost or neighbour maps back through the opposite edge
            nb = c2c[j, c]
            assert c2c[(j + 2) % 6 + 1, nb] == c
    # surface-layer edges come first and map to themselves
    e = np.arange(1, g.szeS)
    assert np.array_equal(U.view(g.m2de, case.sze, np.int32)[e], e)


def test_hex_set_dz_column_walk():
    case = MS.hex_basin(nx=14, ny=12, nz=7, ntr=1, seed=5, dmin=2.0, dmax=40.0)
    rng = np.random.default_rng(1)
    g, wd, wc = case.g, case.wd, case.wc
    eta, old = case.a2(wd.eta), case.a2(wc.oldeta)
    lay = case.layers
    dl = lay[1] - lay[0]
    w2 = U.view(g.w2_t, case.szcS, np.int32)[1:g.v2_t + 1]
    old[w2] = rng.uniform(-1.6, 0.0, len(w2)) * dl
    eta[w2] = old[w2]
    OU.set_dz(case.gp, case.wp, case.pp)
    eta[w2] = old[w2] + rng.uniform(-1.0, 1.0, len(w2)) * dl
    OU.set_dz(case.gp, case.wp, case.pp)
    botz = case.a2(g.botz)
    nz = case.nz
    s1 = U.view(wc.s1, case.szc, np.int32)
    surf, sub = [], []
    dz = case.ac(wc.dz)
    for cs in w2:
        m = case.s2i[cs]
        kb = case.kbot[m]
        kold, knew = nz - 1, nz - 1
        while kold >= kb and lay[kold] >= old[cs]:
            kold -= 1
        while knew >= kb and lay[knew] >= eta[cs]:
            knew -= 1
        klow = min(kold, knew)
        if min(eta[cs], old[cs]) >= 0.05 * wc.hmin + botz[cs] and klow >= kb:
            surf.append(case.idx3[klow, m])
            sub += [case.idx3[k, m] for k in range(klow - 1, kb - 1, -1)]
        for k in range(kb, max(kold, kb - 1) + 1):
            want = min(old[cs], lay[k + 1]) - max(botz[cs], lay[k])
            assert abs(dz[case.idx3[k, m]] - want) < 1e-12
    assert wc.vcs == len(surf)
    assert list(s1[1:wc.vcs + 1]) == surf
    assert list(s1[wc.vcs + 1:wc.vc + 1]) == sub


def test_hex_substep_limit_reports_a_column():
    case = MS.hex_basin(nx=12, ny=10, nz=5, ntr=1, seed=2, dt=1e6, vmax=2.0)
    case.wc.stab = U.SUB_STEP
    case.prepare(OU.set_dz)
    r = OU.advect_diffuse(case)
    assert r > 0
    assert r < case.szcS
