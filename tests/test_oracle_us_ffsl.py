"""Pinning of the COMPAS FFSL path of the oracle (ffsl_don with the
linear least-squares interpolant, oracle/oracle_us.c).

What fixes each part, independently of the formulas restated there:
  streamlines + least squares + divergence
      a linear field carried by a uniform flow over one layer moves
      exactly (x - U dt) in every cell whose stencils see no coast:
      the streamline mean of a linear field is its midpoint value, the
      fit reproduces linear fields, and the edge sum is the divergence
      theorem for a linear integrand
  vertical maps and integer-part integration
      a linear profile under uniform w moves exactly (z - w dt) at
      Courant numbers below and above one
  multi-segment streamlines (hd_trans_interp)
      the velocity fit's points are the polygon neighbourhood of the
      centre and its weights reproduce linear fields; a linear field
      carried two to three cells per step moves exactly
  get_posc / intersect
      end points against polygon geometry computed here: the full
      back-trajectory when it stays in the cell, the crossing with the
      coast when it leaves through a land edge, the next cell otherwise
  get_linear_limit
      the limited reconstruction stays within the range of the centres
      around every vertex (vertex neighbourhoods taken from the mesh
      polygons, not from the maps), and beta is the largest such factor
  the whole step
      constancy and closed-basin mass conservation on hexagons with
      vertical motion; ffsl_init runs on every other call
  clip_ffsl (FILL_METHOD CLIP)
      a translated linear field is left alone in the interior; where the
      clip acts the new value is a nearby cell's step-start value or
      zero; the global range holds"""
import numpy as np
import pytest

from ems_amd import mesh as MS, structs_us as U
from oracle import oracle_us as OU


def _centres(case):
    g = case.g
    m2d = U.view(g.m2d, case.szc, np.int32)
    cs = m2d[case.wet]
    return (U.view(g.cellx, case.szcS, np.float64)[cs],
            U.view(g.celly, case.szcS, np.float64)[cs])


def _flat(nx=24, ny=22, nz=1, seed=5):
    return MS.hex_basin(nx=nx, ny=ny, nz=nz, ntr=1, seed=seed, dmin=10.0 * nz,
                        dmax=10.0 * nz, land_frac=0.0)


def _interior(cx, cy, margin):
    return ((cx > cx.min() + margin) & (cx < cx.max() - margin) &
            (cy > cy.min() + margin) & (cy < cy.max() - margin))


@pytest.mark.parametrize("U0,V0", [(0.6, -0.35), (-0.9, 0.2), (0.05, 0.8)])
def test_linear_field_translates_exactly(U0, V0):
    dt = 180.0
    case = _flat()
    case.set_uniform_velocity(U0, V0)
    case.scheme(U.FFSL, 0, dt)
    cx, cy = _centres(case)
    a, bx, by = 30.0, 2e-4, -3e-4
    t0 = a + bx * cx + by * cy
    case.set_tracers(t0[None, :])
    OU.set_dz(case.gp, case.wp, case.pp)
    assert OU.advect_diffuse(case) == 0
    got = case.tracers()[0]
    exp = a + bx * (cx - U0 * dt) + by * (cy - V0 * dt)
    inner = _interior(cx, cy, 12000.0)
    assert inner.sum() > 50
    assert np.abs(exp - t0)[inner].min() > 1e-3
    assert np.abs(got - exp)[inner].max() < 1e-11


@pytest.mark.parametrize("W0", [0.004, -0.004, 0.025, -0.025])
def test_linear_profile_moves_exactly_vertically(W0):
    """dt * |w| / dz = 0.36 and 2.25 layers."""
    dt, nz = 90.0, 14
    case = _flat(nx=10, ny=9, nz=nz)
    case.set_uniform_velocity(0.0, 0.0)
    case.scheme(U.FFSL, 0, dt)
    g = case.g
    zc = U.view(g.cellz, case.szc, np.float64)
    zm1 = U.view(g.zm1, case.szc, np.int32)
    w = case.ac(case.wd.w)
    bed = zm1[zm1[case.wet]] == zm1[case.wet]
    w[case.wet] = np.where(bed, 0.0, W0)
    t0 = 60.0 + 0.3 * zc[case.wet]
    case.set_tracers(t0[None, :])
    OU.set_dz(case.gp, case.wp, case.pp)
    assert OU.advect_diffuse(case) == 0
    got = case.tracers()[0]
    k = case.s2k[case.wet]
    mid = (k >= 4) & (k <= nz - 5)
    assert np.abs(got - (t0 - 0.3 * W0 * dt))[mid].max() < 1e-11
    assert np.abs(got - t0)[mid].min() > 0.1 * abs(0.3 * W0 * dt)


def _polygon(case, ci):
    mesh = case.mesh
    m = case.s2i[ci]
    v = mesh.vid[m, :mesh.npe[m]]
    return mesh.vx[v], mesh.vy[v]


def _seg_cross(px, py, qx, qy, ax, ay, bx, by):
    d = (qx - px) * (by - ay) - (qy - py) * (bx - ax)
    if d == 0.0:
        return None
    s = ((ax - px) * (by - ay) - (ay - py) * (bx - ax)) / d
    t = ((ax - px) * (qy - py) - (ay - py) * (qx - px)) / d
    if 0.0 <= s <= 1.0 and 0.0 <= t <= 1.0:
        return s
    return None


def test_streamline_end_points():
    case = MS.hex_basin(nx=20, ny=18, nz=3, ntr=1, seed=9)
    g = case.g
    mesh = case.mesh
    rng = np.random.default_rng(2)
    e2c = case.e2c
    m2de = U.view(g.m2de, case.sze, np.int32)
    w3e = U.view(g.w3_e1, case.sze, np.int32)[1:g.a3_e1 + 1]
    wet = set(case.wet.tolist())
    seen = {"inside": 0, "coast": 0, "next": 0}
    for e in rng.choice(w3e, 600, replace=False):
        es = m2de[e]
        mx, my = case.ae2(g.u1x)[es], case.ae2(g.u1y)[es]
        for _ in range(2):
            ang = rng.uniform(0, 2 * np.pi)
            sp = rng.uniform(0.5, 2.0)
            u, v = sp * np.cos(ang), sp * np.sin(ang)
            n = np.array([case.ae2(g.costhu1)[es], case.ae2(g.sinthu1)[es]])
            un = u * n[0] + v * n[1]
            ci = e2c[e, 1] if un > 0 else e2c[e, 0]
            dt = 600.0
            r, x, y, dist = OU.ffsl_posc(case, e, ci, u, v, dt)
            px, py = mx - u * dt, my - v * dt
            vx, vy = _polygon(case, ci)
            m = case.s2i[ci]
            k = case.s2k[ci]
            # first crossing of M->P with an edge of the cell other than e
            best = None
            for j in range(mesh.npe[m]):
                a, b = j, (j + 1) % mesh.npe[m]
                ax, ay, bx_, by_ = vx[a], vy[a], vx[b], vy[b]
                if np.hypot(0.5 * (ax + bx_) - mx, 0.5 * (ay + by_) - my) < 1e-6:
                    continue  # the start edge
                s = _seg_cross(mx, my, px, py, ax, ay, bx_, by_)
                if s is not None and (best is None or s < best[0]):
                    best = (s, j)
            if best is None:
                seen["inside"] += 1
                assert r == -ci
                assert abs(x - px) < 1e-6 and abs(y - py) < 1e-6
                assert abs(dist - np.hypot(u, v) * dt) < 1e-6
                continue
            s, j = best
            ix, iy = mx + s * (px - mx), my + s * (py - my)
            q = mesh.nbr[m, j]
            across_wet = q >= 0 and case.kbot[q] <= k
            assert abs(x - ix) < 1e-6 and abs(y - iy) < 1e-6, (e, ci, x, ix, y, iy)
            assert abs(dist - np.hypot(ix - mx, iy - my)) < 1e-6
            if across_wet:
                seen["next"] += 1
                assert r > 0 and r in wet and case.s2i[r] == q and case.s2k[r] == k
            else:
                seen["coast"] += 1
                assert r == -ci
    assert min(seen.values()) > 5, seen


def test_limited_reconstruction_stays_in_vertex_ranges():
    case = MS.hex_basin(nx=18, ny=16, nz=5, ntr=1, seed=4)
    OU.set_dz(case.gp, case.wp, case.pp)
    g = case.g
    mesh = case.mesh
    rng = np.random.default_rng(8)
    tr = np.zeros(case.szc)
    tr[case.wet] = np.exp(rng.normal(0.0, 1.0, case.nwet))
    # no-gradient ghosts and sediment as ffsl_don leaves them
    bpt = U.view(g.bpt, g.nbpt + 1, np.int32)[1:]
    bin_ = U.view(g.bin, g.nbpt + 1, np.int32)[1:]
    tr[bpt] = tr[bin_]
    W, beta = OU.ffsl_limit(case, tr)
    cellx = U.view(g.cellx, case.szcS, np.float64)
    celly = U.view(g.celly, case.szcS, np.float64)
    m2d = U.view(g.m2d, case.szc, np.int32)
    # centres around each mesh vertex, from the polygons
    around = {}
    for m in range(mesh.nmc):
        for j in range(mesh.npe[m]):
            around.setdefault(int(mesh.vid[m, j]), []).append(m)
    idx3 = case.idx3
    limited = 0
    for c in case.wet:
        m, k, cs = case.s2i[c], case.s2k[c], m2d[c]
        w0, w1, w2 = W[c]
        assert w0 == tr[c]
        b = beta[c]
        assert 0.0 <= b <= 1.0
        slack = []
        for j in range(mesh.npe[m]):
            vv = int(mesh.vid[m, j])
            cells = [idx3[k, q] for q in around[vv] if case.kbot[q] <= k]
            lo, hi = tr[cells].min(), tr[cells].max()
            d = w1 * (cellx[cs] - mesh.vx[vv]) + w2 * (celly[cs] - mesh.vy[vv])
            val = w0 + b * d
            tol = 1e-12 * max(abs(lo), abs(hi), 1.0)
            if d != 0.0 and (w0 + d > hi or w0 + d < lo):
                assert lo - tol <= val <= hi + tol, (c, j, val, lo, hi)
                slack.append(min(abs(val - lo), abs(val - hi)))
        if b < 1.0:
            limited += 1
            assert min(slack) <= 1e-9 * max(abs(tr[c]), 1.0)
    assert limited > 20


def _basin(seed=7):
    case = MS.hex_basin(nx=22, ny=18, nz=7, ntr=3, seed=seed, dt=120.0)
    case.scheme(U.FFSL, 0, 120.0)
    case.prepare(OU.set_dz)
    return case


def test_constancy_and_mass():
    case = _basin()
    t = case.tracers()
    t[0] = 4.25
    case.set_tracers(t)
    m0 = [OU.tmass(case, n) for n in range(case.ntr)]
    for _ in range(4):
        assert OU.step(case) == 0
    got = case.tracers()
    assert np.abs(got[0] - 4.25).max() < 1e-12
    m1 = [OU.tmass(case, n) for n in range(case.ntr)]
    assert np.abs(got[1:] - t[1:]).max() > 1e-3
    for a, b in zip(m0, m1):
        assert abs(b - a) / abs(a) < 1e-12


def test_ffsl_init_every_other_call():
    case = _basin(seed=3)
    g = case.g
    bpt = U.view(g.bpt, g.nbpt + 1, np.int32)[1:]
    u = case.ac(case.wd.u)
    u[bpt] = 123.0
    flags = []
    for _ in range(3):
        assert OU.step(case) == 0
        flags.append(case.wc.flag & U.F_FFSL)
        if len(flags) == 1:
            assert np.all(u[bpt] != 123.0)
            u[bpt] = 123.0
    assert flags == [U.F_FFSL, 0, U.F_FFSL]


# ---- multi-segment streamlines (hd_trans_interp, transport.c:5171-5296) ----

def _vlsq_entries(case, c):
    t = case.vlsq()
    o, n = t.vo[c], t.vn[c]
    return ([t.cell[i] for i in range(o, o + n)], [t.edge[i] for i in range(o, o + n)],
            np.array([t.bx[i] for i in range(o, o + n)]),
            np.array([t.by[i] for i in range(o, o + n)]))


def test_velocity_fit_weights():
    """The streamline velocity fit of every wet centre (the library's
    table, built from the mesh): its wet points are the centre and the
    wet centres of its layer whose polygons share a corner with it (from
    the mesh polygons), its ghost points the ghosts across their land
    edges, placed at the edge midpoints; the slope rows give back the
    slopes of any linear field sampled at those points."""
    case = MS.hex_basin(nx=16, ny=14, nz=4, ntr=1, seed=3, trasc=U.FFSL)
    g, mesh = case.g, case.mesh
    m2d = U.view(g.m2d, case.szc, np.int32)
    m2de = U.view(g.m2de, case.sze, np.int32)
    wgst = U.view(g.wgst, case.szc, np.int32)
    cxs, cys = case.a2(g.cellx), case.a2(g.celly)
    ux, uy = case.ae2(g.u1x), case.ae2(g.u1y)
    corners = [set(mesh.vid[m, :mesh.npe[m]].tolist()) for m in range(mesh.nmc)]
    rng = np.random.default_rng(4)
    nghost = 0
    for c in case.wet:
        m, k = case.s2i[c], case.s2k[c]
        cells, edges, bx, by = _vlsq_entries(case, c)
        wet = [x for x in cells if not wgst[x]]
        ghosts = [(x, e) for x, e in zip(cells, edges) if wgst[x]]
        assert wet[0] == c and len(wet) == len(set(wet))
        want = {case.idx3[k, m2] for m2 in range(mesh.nmc)
                if k >= case.kbot[m2] and corners[m2] & corners[m]}
        assert set(wet) == want, c
        for x, e in ghosts:
            a, b = case.e2c[e]
            assert {a, b} == {x, wgst[x]} and wgst[x] in want
        nghost += len(ghosts)
        px = np.array([cxs[m2d[x]] if not wgst[x] else ux[m2de[e]]
                       for x, e in zip(cells, edges)]) - cxs[m2d[c]]
        py = np.array([cys[m2d[x]] if not wgst[x] else uy[m2de[e]]
                       for x, e in zip(cells, edges)]) - cys[m2d[c]]
        sx, sy = rng.normal(size=2) * 1e-3
        f = 7.0 + sx * px + sy * py
        assert abs(f @ bx - sx) < 1e-12 * 7.0 / 1000.0 + 1e-15
        assert abs(f @ by - sy) < 1e-12 * 7.0 / 1000.0 + 1e-15
    assert nghost > 100


@pytest.mark.parametrize("U0,V0", [(0.9, 0.45), (-0.7, 0.6), (0.15, -1.0)])
def test_multi_segment_linear_field_translates_exactly(U0, V0):
    """A uniform flow crossing two to three cells per step: every
    streamline leaves its first cell and goes on with the fitted
    velocity (uniform, so the path stays straight); the mean of a linear
    field over the segments, each from its own cell's reconstruction, is
    the midpoint value, so the field moves exactly (x - U dt) wherever
    the stencils see no coast."""
    dt = 6000.0
    case = _flat(nx=34, ny=30)
    case.set_uniform_velocity(U0, V0)
    case.scheme(U.FFSL, 0, dt)
    cx, cy = _centres(case)
    a, bx, by = 30.0, 2e-4, -3e-4
    t0 = a + bx * cx + by * cy
    case.set_tracers(t0[None, :])
    OU.set_dz(case.gp, case.wp, case.pp)
    assert OU.advect_diffuse(case) == 0
    got = case.tracers()[0]
    exp = a + bx * (cx - U0 * dt) + by * (cy - V0 * dt)
    inner = _interior(cx, cy, 12000.0 + np.hypot(U0, V0) * dt)
    assert inner.sum() > 40
    assert np.abs(got - exp)[inner].max() < 1e-10
    # the streamlines there really are multi-segment: two cells or more
    assert np.hypot(U0, V0) * dt > 1.6 * 2 * 1500.0 * np.sqrt(3) / 2


def test_multi_segment_needs_a_velocity_fit():
    """Without a fit (ffsl_init has not run on this window: its flag is
    already set on the first call) a multi-segment streamline is
    reported."""
    case = _flat(nx=10, ny=9)
    case.set_uniform_velocity(1.0, 0.5)
    case.scheme(U.FFSL, 0, 3000.0)
    case.set_tracers(np.ones((1, case.nwet)))
    case.wc.flag |= U.F_FFSL
    OU.set_dz(case.gp, case.wp, case.pp)
    assert OU.advect_diffuse(case) == -3


# ---- FILL_METHOD CLIP (clip_ffsl, transport.c:4107-4138) ----------------

def test_clip_leaves_a_translated_linear_field():
    """A linear field in uniform flow moves exactly, and each interior
    cell's new value lies inside the range of the upstream stencils: the
    clip changes nothing there (bounds that were too tight, e.g. only the
    cell's own value, would clip every cell downstream of a gradient)."""
    dt, U0, V0 = 180.0, 0.6, -0.35
    res = {}
    for fill in (0, U.CLIP):
        case = _flat()
        case.set_uniform_velocity(U0, V0)
        case.scheme(U.FFSL, 0, dt)
        case.wc.fillf = fill
        cx, cy = _centres(case)
        t0 = 30.0 + 2e-4 * cx - 3e-4 * cy
        case.set_tracers(t0[None, :])
        OU.set_dz(case.gp, case.wp, case.pp)
        assert OU.advect_diffuse(case) == 0
        res[fill] = case.tracers()[0]
    inner = _interior(cx, cy, 12000.0)
    exp = 30.0 + 2e-4 * (cx - U0 * dt) - 3e-4 * (cy - V0 * dt)
    assert np.abs(res[U.CLIP] - exp)[inner].max() < 1e-11
    assert np.array_equal(res[U.CLIP][inner], res[0][inner])


def _clip_pair(seed, steps=1):
    out = {}
    for fill in (0, U.CLIP):
        case = MS.hex_basin(nx=24, ny=20, nz=8, ntr=3, seed=seed, trasc=U.FFSL, dt=600.0)
        case.wc.fillf = fill
        case.prepare(OU.set_dz)
        t0 = np.stack([case.tr(n).copy() for n in range(case.ntr)])
        for _ in range(steps):
            assert OU.advect_diffuse(case) == 0
        out[fill] = np.stack([case.tr(n).copy() for n in range(case.ntr)])
    return case, t0, out


@pytest.mark.parametrize("seed", [1, 4])
def test_clip_clamps_to_neighbouring_values(seed):
    """Where the clip acts, the new value is exactly the step-start value
    of a cell near the cell (its stencils' cells and its column), or zero
    (a land edge's bounds): the clamp is to data, not to a reconstruction;
    a cell is moved up to a minimum or down to a maximum, never past the
    unclipped value."""
    case, t0, out = _clip_pair(seed)
    g = case.g
    m2d = U.view(g.m2d, case.szc, np.int32)
    cellx = U.view(g.cellx, case.szcS, np.float64)
    celly = U.view(g.celly, case.szcS, np.float64)
    mint = U.view(case.wc.mintr, case.ntr, np.float64)
    wet = case.wet
    x, y = cellx[m2d[wet]], celly[m2d[wet]]
    nclip = 0
    for n in range(case.ntr):
        a, b = out[0][n][wet], out[U.CLIP][n][wet]
        diff = np.nonzero(a != b)[0]
        nclip += diff.size
        for i in diff:
            near = np.hypot(x - x[i], y - y[i]) < 4.0 * 1500.0
            cand = set(t0[n][wet][near].tolist()) | {0.0, float(mint[n])}
            assert float(b[i]) in cand, (n, i, b[i], a[i])
            # clamped toward the range: raised to a minimum or lowered to a
            # maximum, by the amount the unclipped value was out of it
            assert b[i] != a[i]
    assert nclip > 50


def test_clip_keeps_the_global_range():
    case, t0, out = _clip_pair(2, steps=3)
    wet = case.wet
    for n in range(case.ntr):
        lo = min(0.0, t0[n][wet].min())
        hi = max(0.0, t0[n][wet].max())
        got = out[U.CLIP][n][wet]
        assert got.min() >= lo and got.max() <= hi


def test_clip_settings_outside_scope_are_reported():
    for fill, rk in ((U.MONOTONIC, 1), (U.CLIP, 2), (U.CLIP | U.MONOTONIC, 3)):
        case = MS.hex_basin(nx=12, ny=10, nz=4, ntr=1, seed=2, trasc=U.FFSL, dt=300.0)
        case.wc.fillf = fill
        case.wc.rkstage = rk
        case.prepare(OU.set_dz)
        assert OU.advect_diffuse(case) == -6
    # CLIP with MONOTONIC is CLIP (its branch returns before the fill)
    res = []
    for fill in (U.CLIP, U.CLIP | U.MONOTONIC):
        case = MS.hex_basin(nx=12, ny=10, nz=4, ntr=1, seed=2, trasc=U.FFSL, dt=300.0)
        case.wc.fillf = fill
        case.prepare(OU.set_dz)
        assert OU.advect_diffuse(case) == 0
        res.append(case.tracers())
    assert np.array_equal(res[0], res[1])
