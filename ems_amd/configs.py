"""Seeded synthetic windows for the EMS tracer step.

A Case wraps one window built by emstr_grid_build() (libemstr) and fills
the reference's fields: metrics, face volume fluxes and velocities with the
u1flux3d = u1*dzu1*h2au1 convention (model/hd/momentum/vel3d.c:4097),
w from continuity (vel3d.c:4695-4711), Kz on lower faces, elevations and
tracers.  Everything is built from a seed with numpy and is deterministic.

Configurations (sizes are wet-interior sizes; a land ring is added):
  c1  40x40x10 flat 50 m basin, solid-body rotation, 1 Gaussian tracer,
      van Leer, dt = 60 s.
  c2  200x300x30 curvilinear, 12-ridge bathymetry (~30 % land), 16-mode
      streamfunction (max 1 m/s), 10 tracers, QUICKEST, dt = 120 s.
  c3  400x600x40 curvilinear, 25 % land, 2-300 m, max 1.5 m/s, 40 tracers,
      van Leer, dt = 90 s.
  small(...) scaled-down curvilinear cases for tests.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import structs as S


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Case:
    """One window: the C structures plus numpy views of their arrays."""

    def __init__(self, name, nx, ny, nz, layers, botz, dx, dy, ntr):
        self.name = name
        L = _lib.lib()
        self.nx, self.ny, self.nz, self.ntr = nx, ny, nz, ntr
        self.layers = np.ascontiguousarray(layers, dtype=np.float64)
        self.dx = np.asarray(dx, dtype=np.float64)
        self.dy = np.asarray(dy, dtype=np.float64)
        # separable orthogonal metrics at every (i,j) (land included)
        DX, DY = np.meshgrid(self.dx, self.dy)  # [ny, nx]
        dxm = np.concatenate([self.dx[:1], self.dx[:-1]])
        dym = np.concatenate([self.dy[:1], self.dy[:-1]])
        DXM, DYM = np.meshgrid(dxm, dym)
        h1acell = np.ascontiguousarray(DX)
        h2acell = np.ascontiguousarray(DY)
        h1au1 = np.ascontiguousarray(0.5 * (DXM + DX))  # centre-to-centre
        h2au2 = np.ascontiguousarray(0.5 * (DYM + DY))
        h2au1 = np.ascontiguousarray(DY)  # face lengths
        h1au2 = np.ascontiguousarray(DX)
        area = np.ascontiguousarray(DX * DY)
        self.botz2d = np.ascontiguousarray(botz, dtype=np.float64)
        self.metrics = dict(h1acell=h1acell, h2acell=h2acell, h1au1=h1au1,
                            h2au2=h2au2, h2au1=h2au1, h1au2=h1au2,
                            cellarea=area)
        self._keep = [self.layers, self.botz2d, h1acell, h2acell, h1au1,
                      h2au2, h2au1, h1au2, area]
        G = L.emstr_grid_build(nx, ny, nz, _dptr(self.layers),
                               _dptr(self.botz2d), _dptr(h1acell),
                               _dptr(h2acell), _dptr(h1au1), _dptr(h2au2),
                               _dptr(h2au1), _dptr(h1au2), _dptr(area), ntr)
        if not G:
            raise ValueError("emstr_grid_build failed: %s" % (_lib.last_error(),))
        self.G = G
        self.gp = G.contents.window
        self.wp = G.contents.windat
        self.pp = G.contents.wincon
        self.g = self.gp.contents
        self.wd = self.wp.contents
        self.wc = self.pp.contents
        g = self.g
        self.sgsiz, self.sgsizS = g.sgsiz, g.sgsizS
        self.kbot = S.view(G.contents.kbot, nx * ny, np.int32).reshape(ny, nx).copy()
        n = self.sgsiz
        self.s2i = S.view(g.s2i, n, np.int32)
        self.s2j = S.view(g.s2j, n, np.int32)
        self.s2k = S.view(g.s2k, n, np.int32)
        self.nwet = g.b3_t
        self.wet = S.view(g.w3_t, n, np.int32)[1:g.b3_t + 1].copy()
        self.dt = 1.0
        self.steps = 1

    def open_boundary(self, side):
        """Make the wet columns on one side of the grid an open boundary
        (emstr_grid_open); returns its open_bdrys_t.  The wet-cell order
        changes (boundary cells last): call before prepare()."""
        if _lib.lib().emstr_grid_open(self.G, side) != 0:
            raise ValueError("emstr_grid_open failed")
        g = self.g
        self.nwet = g.b3_t
        self.wet = S.view(g.w3_t, self.sgsiz, np.int32)[1:g.b3_t + 1].copy()
        return g.open[g.nobc - 1].contents

    # -- array views -------------------------------------------------
    def a3(self, ptr):
        return S.view(ptr, self.sgsiz, np.float64)

    def a2(self, ptr):
        return S.view(ptr, self.sgsizS, np.float64)

    def i3(self, ptr):
        return S.view(ptr, self.sgsiz, np.int32)

    def i2(self, ptr):
        return S.view(ptr, self.sgsizS, np.int32)

    def tr(self, n):
        return S.view(self.wd.tr_wc[n], self.sgsiz, np.float64)

    def tracers(self):
        """Copy of all tracers over wet cells, [ntr, nwet]."""
        return np.stack([self.tr(n)[self.wet] for n in range(self.ntr)])

    def set_tracers(self, arr):
        for n in range(self.ntr):
            self.tr(n)[self.wet] = arr[n]
        self.fill_ghosts()

    def fill_ghosts(self):
        """Lateral and sediment ghosts take the value of their interior
        cell (tracers.c:5574-5593, 827-830)."""
        g = self.g
        bpt = S.view(g.bpt, g.nbpt + 1, np.int32)[1:]
        bin_ = S.view(g.bin, g.nbpt + 1, np.int32)[1:]
        gs = S.view(g.gsed_t, g.ngsed + 1, np.int32)[1:]
        ist = S.view(g.ised_t, g.ngsed + 1, np.int32)[1:]
        for n in range(self.ntr):
            t = self.tr(n)
            t[bpt] = t[bin_]
            t[gs] = t[ist]

    def dense_index(self, cells):
        """Flat (k,j,i) index of sparse cells."""
        return (self.s2k[cells] * self.ny + self.s2j[cells]) * self.nx + self.s2i[cells]

    def scatter3(self, ptr, dense, cells=None):
        """Write a dense [nz,ny,nx] field into a sparse 3D array at the
        given cells (default: all cells with a layer index in range)."""
        a = self.a3(ptr)
        if cells is None:
            cells = np.arange(1, self.sgsiz, dtype=np.int64)
            cells = cells[self.s2k[cells] >= 0]
        a[cells] = dense.reshape(-1)[self.dense_index(cells)]

    def prepare(self, set_dz_fn):
        """Run set_dz (oracle or GPU) then w from continuity."""
        set_dz_fn(self.gp, self.wp, self.pp)
        _lib.lib().emstr_vel_w_update(self.gp, self.wp, self.pp)

    def free(self):
        if getattr(self, "G", None):
            _lib.lib().emstr_grid_free(self.G)
            self.G = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------
# Generators
# ---------------------------------------------------------------------
def _tanh_spacing(n, rng, lo, hi):
    """Smoothly varying spacing in [lo, hi] from two seeded tanh steps."""
    x = np.arange(n) + 0.5
    s = np.zeros(n)
    for _ in range(2):
        x0 = rng.uniform(0.2, 0.8) * n
        w = rng.uniform(0.05, 0.2) * n
        s += np.tanh((x - x0) / w) * rng.choice([-1.0, 1.0])
    s = (s - s.min()) / max(s.max() - s.min(), 1e-12)
    return lo + (hi - lo) * s


def _smooth_field(X, Y, rng, nmodes=8, kmax=3):
    """Zero-mean, unit-variance smooth random field on the (X,Y) points."""
    Lx = X.max() - X.min() + 1.0
    Ly = Y.max() - Y.min() + 1.0
    f = np.zeros_like(X)
    for _ in range(nmodes):
        kx = rng.integers(1, kmax + 1) * 2 * np.pi / Lx
        ky = rng.integers(1, kmax + 1) * 2 * np.pi / Ly
        f += rng.normal() * np.sin(kx * X + rng.uniform(0, 2 * np.pi)) * \
            np.sin(ky * Y + rng.uniform(0, 2 * np.pi))
    f -= f.mean()
    sd = f.std()
    return f / sd if sd > 0 else f


def _bathymetry(nxi, nyi, Xc, Yc, rng, nridge, land_frac, dmin, dmax):
    """Depth field (positive, m) with nan for land; land_frac of the
    interior is land.  Sum of nridge seeded Gaussian ridges/basins."""
    Lx, Ly = Xc.max(), Yc.max()
    F = np.zeros_like(Xc)
    for _ in range(nridge):
        cx, cy = rng.uniform(0, Lx), rng.uniform(0, Ly)
        sx = rng.uniform(0.05, 0.25) * Lx
        sy = rng.uniform(0.05, 0.25) * Ly
        a = rng.uniform(0.5, 1.0) * rng.choice([-1.0, 1.0])
        F += a * np.exp(-0.5 * (((Xc - cx) / sx) ** 2 + ((Yc - cy) / sy) ** 2))
    thr = np.quantile(F, land_frac)
    top = F.max()
    depth = dmin + (dmax - dmin) * (F - thr) / max(top - thr, 1e-12)
    depth = np.clip(depth, dmin, dmax)
    depth[F < thr] = np.nan
    return depth


def _snap_bed(botz, layers, minfrac=0.25):
    """Make the bed cell at least minfrac of its layer thick."""
    b = botz.copy()
    wet = np.isfinite(b)
    k = np.clip(np.searchsorted(layers, b, side="right") - 1, 0, len(layers) - 2)
    thick = layers[k + 1] - layers[k]
    frac = (layers[k + 1] - b) / thick
    snap = wet & (frac < minfrac)
    b[snap] = layers[k[snap]]
    return b


def _face_fluxes(case, Psi, alpha, vmax):
    """Layer volume fluxes from a corner transport streamfunction.

    Psi [ny+1, nx+1] at cell corners (corner (i,j) is the SW corner of
    cell (i,j)).  The face transport Psi(end) - Psi(start) is shared among
    the open layers of the face in proportion to their thickness, with a
    zero-mean vertical shear of relative size alpha, then scaled so the
    largest |u| is vmax.  Returns dense u1, u2, u1flux3d, u2flux3d, dzu1,
    dzu2 arrays [nz, ny, nx] (west / south faces of each cell).
    """
    nx, ny, nz = case.nx, case.ny, case.nz
    lay = case.layers
    bz = case.botz2d.copy()
    land = case.kbot >= nz
    bz[land] = np.inf
    eta = 0.0
    out = {}
    for d in ("1", "2"):
        if d == "1":
            U = -(Psi[1:, :-1] - Psi[:-1, :-1])  # west face of (i,j): corners (i,j)->(i,j+1)
            bnb = np.concatenate([np.full((ny, 1), np.inf), bz[:, :-1]], axis=1)
            hface = case.metrics["h2au1"]
        else:
            U = Psi[:-1, 1:] - Psi[:-1, :-1]  # south face: corners (i,j)->(i+1,j)
            bnb = np.concatenate([np.full((1, nx), np.inf), bz[:-1, :]], axis=0)
            hface = case.metrics["h1au2"]
        fbed = np.maximum(bz, bnb)  # the shallower bed closes the face
        dzf = np.zeros((nz, ny, nx))
        zc = np.zeros((nz, ny, nx))
        for k in range(nz):
            top = min(lay[k + 1], eta)
            bot = np.maximum(lay[k], fbed)
            th = np.where(np.isfinite(fbed), np.maximum(top - bot, 0.0), 0.0)
            dzf[k] = th
            zc[k] = np.where(th > 0, 0.5 * (top + bot), 0.0)
        H = dzf.sum(axis=0)
        Hs = np.where(H > 0, H, 1.0)
        zbar = (dzf * zc).sum(axis=0) / Hs
        shear = np.where(dzf > 0, 2.0 * (zc - zbar) / Hs, 0.0)
        flux = np.where(dzf > 0, U * (dzf / Hs) * (1.0 + alpha * shear), 0.0)
        vel = np.where(dzf > 0, flux / np.where(dzf > 0, dzf * hface, 1.0), 0.0)
        out[d] = (vel, flux, dzf)
    vm = max(np.abs(out["1"][0]).max(), np.abs(out["2"][0]).max())
    sc = vmax / vm if vm > 0 else 1.0
    u1, f1, dz1 = out["1"]
    u2, f2, dz2 = out["2"]
    return u1 * sc, u2 * sc, f1 * sc, f2 * sc, dz1, dz2


def _fill_fields(case, u1, u2, f1, f2, dz1, dz2, Kz):
    wd, wc = case.wd, case.wc
    case.scatter3(wd.u1, u1)
    case.scatter3(wd.u2, u2)
    case.scatter3(wd.u1flux3d, f1)
    case.scatter3(wd.u2flux3d, f2)
    case.scatter3(wd.dzu1, dz1)
    case.scatter3(wd.dzu2, dz2)
    case.scatter3(wd.Kz, Kz)
    case.a2(wd.eta)[:] = 0.0
    case.a2(wd.topz)[:] = 0.0
    case.a2(wc.oldeta)[:] = 0.0
    case.a2(wd.wbot)[:] = 0.0


def _cell_xy(dx, dy):
    xc = np.cumsum(dx) - 0.5 * dx
    yc = np.cumsum(dy) - 0.5 * dy
    return np.meshgrid(xc, yc)


def _corner_xy(dx, dy):
    xf = np.concatenate([[0.0], np.cumsum(dx)])
    yf = np.concatenate([[0.0], np.cumsum(dy)])
    return np.meshgrid(xf, yf)


def _set_scheme(case, trasc, ultimate, dt, mintr, maxtr):
    wc, wd = case.wc, case.wd
    wc.trasc = trasc
    wc.ultimate = int(ultimate)
    wc.stab = S.NONE
    wd.dt = dt
    wd.dttr = dt
    case.dt = dt
    mn = S.view(wc.mintr, case.ntr, np.float64)
    mx = S.view(wc.maxtr, case.ntr, np.float64)
    mn[:] = mintr
    mx[:] = maxtr


def c1():
    """Flat basin, solid-body rotation, one Gaussian blob, van Leer."""
    nxi = nyi = 40
    nz = 10
    nx, ny = nxi + 2, nyi + 2
    dx = np.full(nx, 1000.0)
    dy = np.full(ny, 1000.0)
    layers = np.linspace(-50.0, 0.0, nz + 1)
    botz = np.full((ny, nx), -50.0)
    botz[0, :] = botz[-1, :] = botz[:, 0] = botz[:, -1] = np.nan
    case = Case("c1", nx, ny, nz, layers, botz, dx, dy, 1)
    # corners of the wet interior are at x = (i-1)*1000, i = 1..nx-1
    X, Y = _corner_xy(dx, dy)
    X = X - 1000.0
    Y = Y - 1000.0
    L = nxi * 1000.0
    r2 = (X - 0.5 * L) ** 2 + (Y - 0.5 * L) ** 2
    psi = 0.5 * r2
    psib = 0.5 * (0.5 * L) ** 2
    psi = np.minimum(psi, psib) - psib  # constant (zero) on the walls
    Psi = psi * 50.0  # depth-integrated transport
    u1, u2, f1, f2, dz1, dz2 = _face_fluxes(case, Psi, 0.0, 0.5)
    z = -np.broadcast_to(layers[:-1][:, None, None], (nz, ny, nx))
    Kz = 1e-4 + 1e-2 * np.exp(-z / 10.0)
    _fill_fields(case, u1, u2, f1, f2, dz1, dz2, Kz)
    Xc, Yc = _cell_xy(dx, dy)
    Xc, Yc = Xc - 1000.0, Yc - 1000.0
    blob = np.exp(-((Xc - 0.3 * L) ** 2 + (Yc - 0.5 * L) ** 2) / (2 * 4000.0 ** 2))
    tr = np.broadcast_to(blob, (nz, ny, nx))
    cells = case.wet
    case.tr(0)[cells] = tr.reshape(-1)[case.dense_index(cells)]
    case.fill_ghosts()
    _set_scheme(case, S.VANLEER, 0, 60.0, [0.0], [1e35])
    case.steps = 200
    return case


def curvilinear(name, nxi, nyi, nz, dmin, dmax, land_frac, vmax, ntr,
                trasc, dt, seed, ultimate=0, n_span=0, alpha=0.5,
                kz_max=1e-1, kz_min=1e-5, hmin=None, top_layer=None):
    """Seeded curvilinear basin (spacing 500-2000 m), ridge bathymetry,
    16-mode streamfunction, mixed-layer Kz; tracers: 2 T/S-like then
    log-normal, the last n_span of which span 1e-6..1e2.  top_layer: the
    thickness of the surface layer (uniform layers below it)."""
    rng = np.random.default_rng(seed)
    nx, ny = nxi + 2, nyi + 2
    dx = _tanh_spacing(nx, rng, 500.0, 2000.0)
    dy = _tanh_spacing(ny, rng, 500.0, 2000.0)
    Xc, Yc = _cell_xy(dx, dy)
    depth = _bathymetry(nxi, nyi, Xc[1:-1, 1:-1], Yc[1:-1, 1:-1], rng, 12,
                        land_frac, dmin, dmax)
    if top_layer is None:
        layers = np.linspace(-dmax, 0.0, nz + 1)
    else:
        layers = np.concatenate([np.linspace(-dmax, -top_layer, nz), [0.0]])
    botz = np.full((ny, nx), np.nan)
    botz[1:-1, 1:-1] = -depth
    botz = _snap_bed(botz, layers)
    case = Case(name, nx, ny, nz, layers, botz, dx, dy, ntr)
    if hmin is not None:
        case.wc.hmin = hmin
    # transport streamfunction, zero on land corners, tapered near coasts
    X, Y = _corner_xy(dx, dy)
    psi = _smooth_field(X, Y, rng, nmodes=16, kmax=4)
    landc = np.ones((ny + 1, nx + 1), dtype=bool)
    wetc = case.kbot < nz
    landc[:] = False
    for a in (0, 1):
        for b in (0, 1):
            landc[a:a + ny, b:b + nx] |= ~wetc
    # corner distance to land (4-neighbour BFS, capped)
    dist = np.where(landc, 0.0, np.inf)
    for it in range(1, 6):
        nb = np.full_like(dist, np.inf)
        nb[1:, :] = np.minimum(nb[1:, :], dist[:-1, :])
        nb[:-1, :] = np.minimum(nb[:-1, :], dist[1:, :])
        nb[:, 1:] = np.minimum(nb[:, 1:], dist[:, :-1])
        nb[:, :-1] = np.minimum(nb[:, :-1], dist[:, 1:])
        dist = np.minimum(dist, nb + 1.0)
    t = np.clip(dist / 5.0, 0.0, 1.0)
    taper = t * t * (3.0 - 2.0 * t)
    Psi = psi * taper
    u1, u2, f1, f2, dz1, dz2 = _face_fluxes(case, Psi, alpha, vmax)
    # Kz: surface mixed layer, depth mld(x,y) in 10..50 m
    mld = 30.0 + 20.0 * np.tanh(_smooth_field(Xc, Yc, rng, 6, 2))
    zf = -np.broadcast_to(layers[:-1][:, None, None], (nz, ny, nx))  # face depth
    Kz = kz_min + (kz_max - kz_min) * 0.5 * (1.0 - np.tanh((zf - mld[None]) / 5.0))
    _fill_fields(case, u1, u2, f1, f2, dz1, dz2, Kz)
    # tracers
    zc = np.broadcast_to(0.5 * (layers[:-1] + layers[1:])[:, None, None], (nz, ny, nx))
    cells = case.wet
    di = case.dense_index(cells)
    mintr = np.zeros(ntr)
    maxtr = np.full(ntr, 1e10)
    for n in range(ntr):
        if n < 2 and ntr > 2:
            t0 = rng.uniform(10.0, 30.0)
            grad = rng.uniform(0.02, 0.05) * (1.0 if n == 0 else -0.2)
            f = t0 + grad * zc + 0.5 * _smooth_field(Xc, Yc, rng)[None]
            mintr[n], maxtr[n] = -5.0, 45.0
        else:
            g2 = _smooth_field(Xc, Yc, rng)
            gz = np.cos(np.pi * zc / layers[0] * rng.uniform(0.5, 2.0))
            G = (g2[None] + 0.5 * gz) / np.sqrt(1.25)
            if n >= ntr - n_span:
                u = 0.5 * (1.0 + np.tanh(G))
                f = 10.0 ** (-6.0 + 8.0 * u)
            else:
                f = np.exp(G)
        case.tr(n)[cells] = np.broadcast_to(f, (nz, ny, nx)).reshape(-1)[di]
    case.fill_ghosts()
    _set_scheme(case, trasc, ultimate, dt, mintr, maxtr)
    return case


def c2(seed=2):
    case = curvilinear("c2", 200, 300, 30, 5.0, 200.0, 0.30, 1.0, 10,
                       S.QUICKEST, 120.0, seed, ultimate=0)
    case.steps = 500
    return case


def c5(seed=5, ndom=64, trasc=None):
    """The multi-domain set: ndom windows with nx, ny ~ U[20, 150],
    5-40 layers and 5-20 tracers each, built as c2; every fourth window is
    stiff (0.5 m surface layer, Kz up to 1 m2/s, dt = 600 s, diffusion
    numbers up to ~1e4).  trasc, if given, is every window's scheme (the
    batched form runs van Leer)."""
    rng = np.random.default_rng(seed)
    out = []
    for d in range(ndom):
        nxi, nyi = (int(v) for v in rng.integers(20, 151, 2))
        nz = int(rng.integers(5, 41))
        ntr = int(rng.integers(5, 21))
        stiff = (d % 4 == 3)
        case = curvilinear("c5_%d" % d, nxi, nyi, nz, 5.0, 200.0, 0.30, 1.0, ntr,
                           trasc if trasc is not None else
                           (S.QUICKEST if not stiff else S.VANLEER),
                           600.0 if stiff else 120.0, int(rng.integers(1 << 30)),
                           kz_max=1.0 if stiff else 1e-1,
                           top_layer=0.5 if stiff else None)
        case.stiff = stiff
        out.append(case)
    return out


def c3(seed=3, ntr=40):
    case = curvilinear("c3", 400, 600, 40, 2.0, 300.0, 0.25, 1.5, ntr,
                       S.VANLEER, 90.0, seed, n_span=4)
    case.steps = 1000
    return case


def small(seed=7, nxi=24, nyi=20, nz=8, ntr=3, trasc=S.VANLEER, ultimate=0,
          dt=60.0, dmin=3.0, dmax=60.0, land_frac=0.25, vmax=0.6, n_span=1):
    return curvilinear("small", nxi, nyi, nz, dmin, dmax, land_frac, vmax, ntr,
                       trasc, dt, seed, ultimate=ultimate, n_span=n_span)


CONFIGS = {"c1": c1, "c2": c2, "c3": c3}
