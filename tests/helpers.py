"""Small windows with hand-set fields for closed-form and parity tests."""
import numpy as np

from ems_amd import configs, structs as S


def box(nxi=12, nyi=10, nz=5, depth=25.0, dx=1000.0, dy=1000.0, ntr=1,
        botz=None):
    """Flat (or given botz) closed box with a land ring, zero fields."""
    nx, ny = nxi + 2, nyi + 2
    layers = np.linspace(-depth, 0.0, nz + 1)
    if botz is None:
        b = np.full((ny, nx), -depth)
    else:
        b = np.full((ny, nx), np.nan)
        b[1:-1, 1:-1] = botz
    b[0, :] = b[-1, :] = b[:, 0] = b[:, -1] = np.nan
    case = configs.Case("box", nx, ny, nz, layers, b, np.full(nx, dx),
                        np.full(ny, dy), ntr)
    wd, wc = case.wd, case.wc
    for name in ("eta", "topz", "wbot"):
        case.a2(getattr(wd, name))[:] = 0.0
    case.a2(wc.oldeta)[:] = 0.0
    return case


def set_dense_face_fields(case, u1=None, u2=None, w=None, Kz=None):
    """Set u1/u2 (and fluxes u*dzu*h) from dense [nz,ny,nx] arrays; faces
    touching land are closed."""
    nz, ny, nx = case.nz, case.ny, case.nx
    wet = np.zeros((nz, ny, nx), bool)
    for k in range(nz):
        wet[k] = case.kbot <= k
    lay = case.layers
    dzl = np.diff(lay)[:, None, None] * np.ones((nz, ny, nx))
    h2au1 = case.metrics["h2au1"][None]
    h1au2 = case.metrics["h1au2"][None]
    if u1 is not None:
        open1 = wet & np.roll(wet, 1, axis=2)
        open1[:, :, 0] = False
        U = np.where(open1, u1, 0.0)
        case.scatter3(case.wd.u1, U)
        case.scatter3(case.wd.u1flux3d, U * dzl * h2au1)
        case.scatter3(case.wd.dzu1, np.where(open1, dzl, 0.0))
    if u2 is not None:
        open2 = wet & np.roll(wet, 1, axis=1)
        open2[:, 0, :] = False
        V = np.where(open2, u2, 0.0)
        case.scatter3(case.wd.u2, V)
        case.scatter3(case.wd.u2flux3d, V * dzl * h1au2)
        case.scatter3(case.wd.dzu2, np.where(open2, dzl, 0.0))
    if w is not None:
        case.scatter3(case.wd.w, np.where(wet, w, 0.0))
    if Kz is not None:
        case.scatter3(case.wd.Kz, np.where(wet, Kz, 0.0))


def set_tracer_dense(case, n, arr):
    cells = case.wet
    case.tr(n)[cells] = np.broadcast_to(arr, (case.nz, case.ny, case.nx)).reshape(-1)[
        case.dense_index(cells)]
    case.fill_ghosts()


def tracer_dense(case, n):
    out = np.full((case.nz, case.ny, case.nx), np.nan)
    cells = case.wet
    out.reshape(-1)[case.dense_index(cells)] = case.tr(n)[cells]
    return out


def scheme(case, trasc, ultimate=0, dt=60.0, mintr=0.0, maxtr=1e35):
    wc, wd = case.wc, case.wd
    wc.trasc = trasc
    wc.ultimate = ultimate
    wc.stab = S.NONE
    wd.dt = wd.dttr = dt
    case.dt = dt
    S.view(wc.mintr, case.ntr, np.float64)[:] = mintr
    S.view(wc.maxtr, case.ntr, np.float64)[:] = maxtr


def _lists(case):
    """set_dz's cell list s1 and its bed cells i1 (reset_dz in
    auxiliary_routines writes its own list into i1, tracers.c:5948),
    as numpy views over their sizes (3D and 2D sparse)."""
    wc = case.wc
    return [S.view(wc.s1, case.sgsiz, np.int32), S.view(wc.i1, case.sgsizS, np.int32)]


def snapshot(case):
    """The tracers and, on a z-level (SHOC) window, the layer thicknesses
    and the cell lists s1/i1 (tracer_step_3d's auxiliary_routines resets
    the surface thicknesses and i1, tracers.c:289)."""
    snap = [case.tr(n).copy() for n in range(case.ntr)]
    if hasattr(case, "a3"):
        snap += [case.a3(case.wc.dz).copy()] + [a.copy() for a in _lists(case)]
    return snap


def restore(case, snap):
    for n in range(case.ntr):
        case.tr(n)[:] = snap[n]
    if hasattr(case, "a3"):
        case.a3(case.wc.dz)[:] = snap[case.ntr]
        for a, b in zip(_lists(case), snap[case.ntr + 1:]):
            a[:] = b
