"""Pins of the oracle's open-boundary tracer path (bdry_tracer /
set_OBC_tr, model/hd/tracers/tracers.c:4542-5165, and the exclusion of
boundary cells from the advective update, tracers.c:811-822) by what
each condition means, computed here independently: the boundary cells
are not advected (their values come through the step unchanged when the
condition writes nothing), NOGRAD copies the cell one inwards, CLAMPD
sets the value, LINEXT continues the least-squares line through the four
cells inwards (numpy's polyfit), FILEIN sets the data below trpc and
NOGRAD acts above it, and UPSTRM moves the boundary value towards the
outside value on inflow and towards the interior on outflow by the
Courant number of the boundary velocity (the textbook upwind step; a
wrong sign of the edge or of the velocity swaps the two)."""
import ctypes as C

import numpy as np

from ems_amd import configs, structs as S
from oracle import oracle as O


def obc_case(seed=3, side=S.L_EDGE, ntr=3, dt=60.0):
    case = configs.small(seed=seed, ntr=ntr, nz=6, dt=dt, land_frac=0.1)
    o = case.open_boundary(side)
    case.keep = []
    case.prepare(O.set_dz)
    return case, o


def cells(case, o):
    m = o.no3_t
    obc = S.view(o.obc_t, m + 1, np.int32)[1:].copy()
    oi1 = S.view(o.oi1_t, m + 1, np.int32)[1:].copy()
    return obc, oi1


def set_conditions(o, conds, clamp=None, trpc=None):
    for n, b in enumerate(conds):
        o.bcond_tra[n] = b
        if clamp is not None:
            o.clampv[n] = clamp[n]
        if trpc is not None:
            o.trpc[n] = trpc[n]


def zero_kz(case):
    case.a3(case.wd.Kz)[:] = 0.0


def test_boundary_cells_are_not_advected():
    """NOTHIN everywhere and no vertical mixing: the boundary cells keep
    their values (to the rounding of the identity solve) while the
    interior moves; the cells above each boundary column's surface take
    the surface value."""
    case, o = obc_case(3)
    zero_kz(case)
    set_conditions(o, [S.NOTHIN] * case.ntr)
    obc, oi1 = cells(case, o)
    t0 = [case.tr(n).copy() for n in range(case.ntr)]
    assert O.step(case) == 0
    g = case.g
    zp1 = case.i3(g.zp1)
    for n in range(case.ntr):
        t = case.tr(n)
        assert np.allclose(t[obc], t0[n][obc], rtol=1e-14, atol=0)
        assert np.abs(t[case.wet] - t0[n][case.wet]).max() > 1e-6
    assert case.g.v2_t < case.g.b2_t and case.g.v3_t < case.g.b3_t


def test_nograd_clampd_linext():
    case, o = obc_case(4)
    set_conditions(o, [S.NOGRAD, S.CLAMPD, S.LINEXT], clamp=[0.0, 7.5, 0.0])
    obc, oi1 = cells(case, o)
    assert O.step(case) == 0
    t0, t1, t2 = case.tr(0), case.tr(1), case.tr(2)
    assert np.array_equal(t0[obc], t0[oi1])
    assert np.all(t1[obc] == 7.5)
    nmap = case.i3(o.nmap)
    for c in obc:
        chain = [nmap[c]]
        for _ in range(3):
            chain.append(nmap[chain[-1]])
        y = t2[chain[::-1]]  # x = 1 .. 4 from the fourth cell inwards to the first
        a, b = np.polyfit([1.0, 2.0, 3.0, 4.0], y, 1)
        want = a * 5.0 + b
        assert abs(t2[c] - want) <= 1e-12 * max(1.0, abs(want)), c


def test_filein_below_trpc_nograd_above():
    case, o = obc_case(5)
    m = o.no3_t
    rng = np.random.default_rng(1)
    data = np.concatenate([[0.0], rng.uniform(10.0, 20.0, m)])
    rows = (S._PD * 1)()
    rows[0] = data.ctypes.data_as(S._PD)
    case.keep += [data, rows]
    o.t_transfer = rows
    o.trm[0] = 0
    depth = -8.0
    set_conditions(o, [S.NOGRAD | S.FILEIN, S.NOGRAD, S.NOGRAD], trpc=[depth, 0.0, 0.0])
    obc, oi1 = cells(case, o)
    assert O.step(case) == 0
    t = case.tr(0)
    cz = case.a3(case.g.cellz)[obc]
    below = cz <= depth
    assert below.any() and (~below).any()
    assert np.array_equal(t[obc][below], data[1:][below])
    assert np.array_equal(t[obc][~below], t[oi1][~below])


def test_upstream_direction():
    """UPSTRM on a west boundary with the boundary face velocity (FACE)
    and CLAMPD as the outside value: against the same step with NOTHIN
    (which leaves the boundary cells and the interior as they are before
    the boundary acts)."""
    vals = {}
    for conds in ([S.NOTHIN] * 3, [S.UPSTRM | S.CLAMPD] * 3):
        case, o = obc_case(6)
        zero_kz(case)
        o.upmeth = S.FACE
        set_conditions(o, conds, clamp=[2.0, 3.0, 4.0])
        obc, oi1 = cells(case, o)
        u1 = case.a3(case.wd.u1)
        rng = np.random.default_rng(2)
        fb = S.view(o.obc_e1, o.no3_t + 1, np.int32)[1:]
        u1[fb] = rng.uniform(-0.5, 0.5, fb.size)
        assert O.step(case) == 0
        vals[conds[0]] = [case.tr(n).copy() for n in range(3)]
    g = case.g
    h = case.a2(g.h1au1)[case.i3(g.m2d)[obc]]
    v = u1[fb]
    cr = 60.0 * np.abs(v) / h
    for n, clamp in enumerate([2.0, 3.0, 4.0]):
        pre = vals[S.NOTHIN][n]
        got = vals[S.UPSTRM | S.CLAMPD][n][obc]
        tb, ti = pre[obc], pre[oi1]
        # inflow (into the domain: u1 > 0 at a west face) brings the
        # outside value, outflow carries the interior value out
        want = np.where(v > 0, tb + cr * (clamp - tb), tb - cr * (tb - ti))
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13)
        assert (v > 0).any() and (v < 0).any()


def test_unsupported_conditions_reported():
    case, o = obc_case(7)
    set_conditions(o, [S.NOGRAD, S.CYCLIC, S.NOGRAD])
    assert O.step(case) == 9
