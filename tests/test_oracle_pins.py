"""Pins the nonlinear terms of the CPU oracles on inputs where each term
is active, against values computed here from the published definitions
of the schemes (not from the oracle's code):

* QUICKEST (Leonard 1979) and its uniform-grid form QUICKEST_CO: with a
  uniform grid and velocity the scheme is exact for cubic profiles, so a
  cubic translates exactly, x - U dt.  The curvature term
  (1 - c^2) curv / 6 (tracers.c:1866-1884) is what makes it exact: on a
  cubic its face-to-face difference is non-zero, so a wrong coefficient,
  sign or upwind choice fails (a quadratic would not pin it: there the
  curvature is the same at every face and cancels in the divergence);
* ORDER4 (tracers.c:1741-1806): its face value (7 (a + b) - (c + d)) / 12
  is exact for cell averages of a cubic, so one step equals the closed
  form  mean_i - U dt (f(x_i + h/2) - f(x_i - h/2)) / h;
* van Leer (1977): face value  t_u + (1 - |c|) phi(r) (t_d - t_u) / 2  with
  phi(r) = (r + |r|) / (1 + |r|),  r = (t_u - t_uu) / (t_d - t_u),  on
  random (curved, non-monotone, extremal) profiles; minmod or superbee in
  its place fail;
* ULTIMATE (Leonard 1991) over QUICKEST on random profiles: the
  normalised-variable bounds, first-order upwind outside 0 <= n_c <= 1.
  Its lower bound n_f < n_c cannot act here: on a uniform grid QUICKEST
  gives n_f >= n_c for every 0 <= n_c <= 1 (n_f - n_c is linear in n_c,
  (1 - c)(2 - c) / 6 at 0 and (1 - c^2) / 6 at 1);
* surf_conc (tracers.c:3047-3207) when the surface rises or falls across
  two layers: the new surface value is the tracer content of the water
  from the processed cell's bottom face to the old surface, plus the
  inflow through that face, minus the horizontal divergence of every
  layer from it to the top, over the new surface volume; computed here
  on dense (k, j, i) arrays with first-order upwind faces.

Each check runs through the SHOC oracle (oracle.c) along x, y and z, and
through the COMPAS oracle (oracle_us.c) on the quadrilateral mesh of the
same window, where its edge stencils (c2c/e2e), signs and spacings must
give the same values."""
import numpy as np
import pytest

from ems_amd import _lib, mesh as MS, structs as S, structs_us as U
from oracle import oracle as O, oracle_us as OU
from tests import helpers as H

H_AX = (1000.0, 1000.0, 5.0)  # cell size along x, y (m) and layer thickness


def _line(axis, vel, scheme, dt, n=20):
    """A closed box with a long axis (0: x, 1: y, 2: z) and a uniform
    velocity along it; zero diffusion, no positivity clip."""
    if axis == 0:
        case = H.box(nxi=n, nyi=5, nz=3, depth=15.0)
    elif axis == 1:
        case = H.box(nxi=5, nyi=n, nz=3, depth=15.0)
    else:
        case = H.box(nxi=3, nyi=3, nz=n, depth=H_AX[2] * n)
    nz, ny, nx = case.nz, case.ny, case.nx
    zero = np.zeros((nz, ny, nx))
    full = np.full((nz, ny, nx), vel)
    H.set_dense_face_fields(case, u1=full if axis == 0 else zero,
                            u2=full if axis == 1 else zero, w=zero, Kz=zero)
    H.scheme(case, scheme, 0, dt, mintr=-1e10)
    case.prepare(O.set_dz)
    # prepare() sets w from continuity; the box is divergent at its ends
    case.a3(case.wd.w)[:] = vel if axis == 2 else 0.0
    return case


def _coord(case, axis):
    """Cell-centre position along the axis, in cell sizes, as a dense
    [nz, ny, nx] array."""
    nz, ny, nx = case.nz, case.ny, case.nx
    if axis == 0:
        s = (np.arange(nx) + 0.5)[None, None, :]
    elif axis == 1:
        s = (np.arange(ny) + 0.5)[None, :, None]
    else:
        zc = 0.5 * (case.layers[:-1] + case.layers[1:]) / H_AX[2]
        s = zc[:, None, None]
    return np.broadcast_to(s, (nz, ny, nx)).copy()


def _inner(case, axis, a):
    """Cells at least four cells from the ends of the long axis (the
    stencil of one step reaches two)."""
    if axis == 0:
        return a[:, 1:-1, 4:case.nx - 4]
    if axis == 1:
        return a[:, 4:case.ny - 4, 1:-1]
    return a[3:case.nz - 3, 1:-1, 1:-1]


def _step_shoc(case):
    assert O.advect_diffuse(case) == 0
    return H.tracer_dense(case, 0)


def _step_quads(case):
    """The same advection step through the COMPAS oracle on the
    quadrilateral mesh of the window; the result on the dense grid."""
    qc = MS.quads_from_case(case)
    OU.set_dz(qc.gp, qc.wp, qc.pp)
    OU.set_dz(qc.gp, qc.wp, qc.pp)
    assert OU.advect_diffuse(qc) == 0
    out = np.full((case.nz, case.ny, case.nx), np.nan)
    kk, jj, ii = qc.shoc_dense
    out[kk, jj, ii] = qc.tracers()[0]
    return out


STEP = {"shoc": _step_shoc, "compas": _step_quads}


def _cubic(xi):
    return 5.0 + 0.3 * xi - 0.02 * xi ** 2 + 0.004 * xi ** 3


def _cubic_int(xi):
    return 5.0 * xi + 0.15 * xi ** 2 - 0.02 / 3.0 * xi ** 3 + 0.001 * xi ** 4


# ---------------------------------------------------------------------
# QUICKEST / QUICKEST_CO: exact for cubics
# ---------------------------------------------------------------------
@pytest.mark.parametrize("which,scheme", [("shoc", S.QUICKEST), ("shoc", S.QUICKEST_CO),
                                          ("compas", S.QUICKEST)])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cour", [0.4, -0.3])
def test_quickest_translates_cubic_exactly(which, scheme, axis, cour):
    if which == "compas" and axis == 2:
        pytest.skip("the COMPAS vertical fluxes are the SHOC column code (checked with shoc)")
    dt = 1000.0
    vel = cour * H_AX[axis] / dt
    case = _line(axis, vel, scheme, dt)
    s = _coord(case, axis)
    xi0 = s[(0, 0, 0)] * 0 + (s.max() + s.min()) / 2
    H.set_tracer_dense(case, 0, _cubic(s - xi0))
    got = _inner(case, axis, STEP[which](case))
    want = _inner(case, axis, _cubic(s - xi0 - cour))
    assert np.abs(got - want).max() < 1e-11
    # the curvature term is active: Lax-Wendroff alone (no curvature) is
    # off by c (1 - c^2) t''' h^3 / 6 on a cubic
    assert abs(cour * (1 - cour ** 2) * 6 * 0.004 / 6) > 1e-4


# ---------------------------------------------------------------------
# ORDER4: exact face values for cell averages of a cubic
# ---------------------------------------------------------------------
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cour", [0.4, -0.3])
def test_order4_cubic_cell_means(axis, cour):
    dt = 1000.0
    vel = cour * H_AX[axis] / dt
    case = _line(axis, vel, S.ORDER4, dt)
    s = _coord(case, axis)
    xi = s - (s.max() + s.min()) / 2
    mean = _cubic_int(xi + 0.5) - _cubic_int(xi - 0.5)
    H.set_tracer_dense(case, 0, mean)
    got = _inner(case, axis, _step_shoc(case))
    want = _inner(case, axis, mean - cour * (_cubic(xi + 0.5) - _cubic(xi - 0.5)))
    assert np.abs(got - want).max() < 1e-11


# ---------------------------------------------------------------------
# van Leer: the harmonic limiter on random profiles
# ---------------------------------------------------------------------
def _phi_vanleer(r):
    return (r + np.abs(r)) / (1.0 + np.abs(r))


def _phi_minmod(r):
    return np.maximum(0.0, np.minimum(1.0, r))


def _limited_step(t, cour, phi):
    """One step of the limited upwind scheme on a 1-D profile t (cell
    values along the axis), velocity sign and Courant number cour; faces
    f = i - 1/2 for i = 2..n-2; returns the new values of cells 2..n-3."""
    n = len(t)
    i = np.arange(2, n - 1)  # faces between cells i - 1 and i
    if cour > 0:
        tu, tuu, td = t[i - 1], t[i - 2], t[i]
    else:
        tu, tuu, td = t[i], t[np.minimum(i + 1, n - 1)], t[i - 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        r = (tu - tuu) / (td - tu)
    F = tu + 0.5 * (1.0 - abs(cour)) * phi(r) * (td - tu)
    # cell j = 2..n-3 has west face index j - 2 and east face j - 1
    return t[2:n - 2] - cour * (F[1:] - F[:-1])


def _random_lines(case, axis, seed):
    """Independent random profiles along the axis in every row (or
    column) of the box, as a dense [nz, ny, nx] array."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (case.nz, case.ny, case.nx))


def _along(fn, a, axis, *args):
    """fn applied to every 1-D line of the dense array a along the axis
    (x: last array axis, z: first)."""
    return np.apply_along_axis(lambda t: fn(t, *args), 2 - axis, a)


def _core(case, axis, a, full):
    """Inner cells of a dense result (full=True) or of a line function's
    output, which starts at cell 2 of each line."""
    n = (case.nx, case.ny, case.nz)[axis]
    lo = 4 if axis < 2 else 3
    idx = [slice(1, -1)] * 3
    idx[2 - axis] = slice(lo, n - lo) if full else slice(lo - 2, n - 2 - lo)
    return a[tuple(idx)]


@pytest.mark.parametrize("which", ["shoc", "compas"])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cour", [0.45, -0.35])
def test_van_leer_limiter_on_random_profiles(which, axis, cour):
    if which == "compas" and axis == 2:
        pytest.skip("the COMPAS vertical fluxes are the SHOC column code (checked with shoc)")
    dt = 1000.0
    vel = cour * H_AX[axis] / dt
    case = _line(axis, vel, S.VANLEER, dt)
    t0 = _random_lines(case, axis, 11 + axis)
    H.set_tracer_dense(case, 0, t0)
    got = _core(case, axis, STEP[which](case), True)
    want = _core(case, axis, _along(_limited_step, t0, axis, cour, _phi_vanleer), False)
    other = _core(case, axis, _along(_limited_step, t0, axis, cour, _phi_minmod), False)
    assert np.abs(got - want).max() < 1e-13
    # the inputs tell the limiters apart
    assert np.abs(want - other).max() > 1e-3


# ---------------------------------------------------------------------
# QUICKEST + ULTIMATE: the normalised-variable bounds on random profiles
# ---------------------------------------------------------------------
def _quickest_ultimate_step(t, cour, ultimate=True):
    """QUICKEST face values on a uniform grid (upstream t_u, downstream
    t_d, far upstream t_uu), then Leonard's (1991) ULTIMATE bounds in
    normalised variables  n_c = (t_u - t_uu) / (t_d - t_uu),
    n_f = (F - t_uu) / (t_d - t_uu):  for 0 <= n_c <= 1 the face is kept in
    [n_c, n_c / |c|] and below 1, with the bound F <= t_d tested first as
    the reference does (tracers.c:3546-3554); otherwise first-order
    upwind.  Returns the new values of cells 2..n-3."""
    n = len(t)
    i = np.arange(2, n - 1)
    if cour > 0:
        tu, tuu, td = t[i - 1], t[i - 2], t[i]
    else:
        tu, tuu, td = t[i], t[np.minimum(i + 1, n - 1)], t[i - 1]
    c = abs(cour)
    F = 0.5 * (tu + td) - 0.5 * c * (td - tu) - (1.0 - c * c) * (td - 2.0 * tu + tuu) / 6.0
    if ultimate:
        d = td - tuu
        nz = d != 0.0
        nc = np.where(nz, (tu - tuu) / np.where(nz, d, 1.0), 0.0)
        nf = np.where(nz, (F - tuu) / np.where(nz, d, 1.0), 0.0)
        inside = (nc >= 0.0) & (nc <= 1.0)
        G = np.where(nf > 1.0, td,
                     np.where(nf < nc, tuu + nc * d,
                              np.where(nf > nc / c, tuu + nc * d / c, F)))
        F = np.where(inside, G, tuu + nc * d)
    return t[2:n - 2] - cour * (F[1:] - F[:-1])


@pytest.mark.parametrize("which", ["shoc", "compas"])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cour", [0.45, -0.35])
def test_quickest_ultimate_on_random_profiles(which, axis, cour):
    if which == "compas" and axis == 2:
        pytest.skip("the COMPAS vertical fluxes are the SHOC column code (checked with shoc)")
    dt = 1000.0
    vel = cour * H_AX[axis] / dt
    case = _line(axis, vel, S.QUICKEST, dt)
    case.wc.ultimate = 1
    t0 = _random_lines(case, axis, 23 + axis)
    H.set_tracer_dense(case, 0, t0)
    got = _core(case, axis, STEP[which](case), True)
    want = _core(case, axis, _along(_quickest_ultimate_step, t0, axis, cour), False)
    plain = _core(case, axis, _along(_quickest_ultimate_step, t0, axis, cour, False), False)
    assert np.abs(got - want).max() < 1e-13
    # the bounds act on these inputs
    assert np.abs(want - plain).max() > 1e-3


# ---------------------------------------------------------------------
# surf_conc: the surface crossing layers
# ---------------------------------------------------------------------
def _moving_box(old, new, seed):
    """6 x 5 columns of six 5 m layers, first-order upwind, random
    horizontal velocities and w from continuity; the old surface at
    `old`, the new one at `new` (per column arrays [nyi, nxi])."""
    case = H.box(nxi=6, nyi=5, nz=6, depth=30.0, ntr=1)
    nz, ny, nx = case.nz, case.ny, case.nx
    rng = np.random.default_rng(seed)
    H.set_dense_face_fields(case, u1=rng.uniform(-0.4, 0.4, (nz, ny, nx)),
                            u2=rng.uniform(-0.4, 0.4, (nz, ny, nx)),
                            w=np.zeros((nz, ny, nx)), Kz=np.zeros((nz, ny, nx)))
    H.scheme(case, S.ORDER1, 0, 300.0, mintr=-1e10)
    g = case.g
    w2 = S.view(g.w2_t, case.sgsizS, np.int32)[1:g.v2_t + 1]
    eta, oe = case.a2(case.wd.eta), case.a2(case.wc.oldeta)
    ij = (case.s2j[w2] - 1, case.s2i[w2] - 1)
    oe[w2] = old[ij]
    eta[w2] = old[ij]
    O.set_dz(case.gp, case.wp, case.pp)  # the previous step ends at old
    eta[w2] = new[ij]
    case.prepare(O.set_dz)
    # tracers: random below the old surface, copied up above it
    t = rng.uniform(1.0, 3.0, (nz, ny, nx))
    lay = case.layers
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            ko = np.searchsorted(lay, old[j - 1, i - 1], side="left") - 1
            t[ko + 1:, j, i] = t[ko, j, i]
    H.set_tracer_dense(case, 0, t)
    return case, t


def _dense(case, ptr):
    d = np.zeros(case.nz * case.ny * case.nx)
    d[case.dense_index(case.wet)] = case.a3(ptr)[case.wet]
    return d.reshape(case.nz, case.ny, case.nx)


def _surface_expected(case, t, old, new):
    """New value of each column's processed surface cell from the tracer
    content balance, with first-order upwind face values."""
    nz, ny, nx = case.nz, case.ny, case.nx
    dt = case.dt
    F1, F2, W = _dense(case, case.wd.u1flux3d), _dense(case, case.wd.u2flux3d), \
        _dense(case, case.wd.w)
    # outflow of every cell through its four faces, tracer units x m3
    Fx = np.zeros((nz, ny, nx + 1))
    Fy = np.zeros((nz, ny + 1, nx))
    for i in range(1, nx):
        up = np.where(F1[:, :, i] > 0, t[:, :, i - 1], t[:, :, i])
        Fx[:, :, i] = up * F1[:, :, i] * dt
    for j in range(1, ny):
        up = np.where(F2[:, j, :] > 0, t[:, j - 1, :], t[:, j, :])
        Fy[:, j, :] = up * F2[:, j, :] * dt
    div = Fx[:, :, 1:] - Fx[:, :, :-1] + Fy[:, 1:, :] - Fy[:, :-1, :]
    lay = case.layers
    area = case.metrics["cellarea"]
    exp = {}
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            kold = np.searchsorted(lay, old[j - 1, i - 1], side="left") - 1
            knew = np.searchsorted(lay, new[j - 1, i - 1], side="left") - 1
            ks = min(kold, knew)
            a = area[j, i]
            # inflow through the bottom face of the processed cell
            wb = W[ks, j, i]
            up = t[ks - 1, j, i] if wb > 0 else t[ks, j, i]
            m = up * wb * dt * a
            for k in range(ks, kold + 1):
                top = lay[k + 1] if k < kold else old[j - 1, i - 1]
                m += t[k, j, i] * a * (top - lay[k])
            m -= div[ks:, j, i].sum()
            exp[(j, i)] = (ks, m / ((new[j - 1, i - 1] - lay[ks]) * a))
    return exp


@pytest.mark.parametrize("which", ["shoc", "compas"])
@pytest.mark.parametrize("move", ["fall", "rise", "mixed"])
def test_surf_conc_across_layers(which, move):
    rng = np.random.default_rng({"fall": 1, "rise": 2, "mixed": 3}[move])
    nyi, nxi = 5, 6
    hi = rng.uniform(-4.5, -0.5, (nyi, nxi))    # in the top layer
    lo = rng.uniform(-14.5, -10.5, (nyi, nxi))  # two layers lower
    if move == "fall":
        old, new = hi, lo
    elif move == "rise":
        old, new = lo, hi
    else:
        old = rng.uniform(-14.5, -0.5, (nyi, nxi))
        new = rng.uniform(-14.5, -0.5, (nyi, nxi))
    case, t = _moving_box(old, new, seed=7)
    exp = _surface_expected(case, t, old, new)
    if which == "shoc":
        got = _step_shoc(case)
    else:
        qc = MS.quads_from_case(case)
        # the COMPAS window settles at the old surface, then moves
        qe, q2 = qc.a2(qc.wd.eta), qc.a2(qc.wc.oldeta)
        cs_q = U.view(qc.g.w2_t, qc.szcS, np.int32)[1:qc.g.v2_t + 1]
        m = qc.s2i[cs_q]
        jj, ii = m // case.nx, m % case.nx
        q2[cs_q] = old[jj - 1, ii - 1]
        qe[cs_q] = old[jj - 1, ii - 1]
        OU.set_dz(qc.gp, qc.wp, qc.pp)
        qe[cs_q] = new[jj - 1, ii - 1]
        OU.set_dz(qc.gp, qc.wp, qc.pp)
        # w as in the structured window (continuity with the same fluxes)
        assert OU.advect_diffuse(qc) == 0
        got = np.full((case.nz, case.ny, case.nx), np.nan)
        kk, jd, idd = qc.shoc_dense
        got[kk, jd, idd] = qc.tracers()[0]
    crossed = 0
    for (j, i), (ks, v) in exp.items():
        assert abs(got[ks, j, i] - v) <= 1e-12 * abs(v), (j, i, ks, got[ks, j, i], v)
        kold = np.searchsorted(case.layers, old[j - 1, i - 1]) - 1
        knew = np.searchsorted(case.layers, new[j - 1, i - 1]) - 1
        crossed += abs(kold - knew) >= 2
    if move != "mixed":
        assert crossed == len(exp)


def test_surf_conc_uses_w_from_the_window():
    """The balance above reads w from the window: make sure it is not zero
    at the processed cells' bottom faces (w from continuity of the random
    horizontal flows), so the inflow term is exercised."""
    rng = np.random.default_rng(1)
    old = rng.uniform(-4.5, -0.5, (5, 6))
    new = rng.uniform(-14.5, -10.5, (5, 6))
    case, t = _moving_box(old, new, seed=7)
    W = _dense(case, case.wd.w)
    ks = np.searchsorted(case.layers, new) - 1
    vals = np.array([W[ks[j, i], j + 1, i + 1] for j in range(5) for i in range(6)])
    assert np.abs(vals).min() > 0.0
    _lib.lib()  # the library the window builder lives in is loaded
