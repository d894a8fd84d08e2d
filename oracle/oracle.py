"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    r = subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"],
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIBPATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        from ems_amd import structs as S
        L = C.CDLL(LIBPATH)
        G, W, P = C.POINTER(S.Geometry), C.POINTER(S.Window), C.POINTER(S.WinPriv)
        for fn in ("oracle_set_dz", "oracle_tracer_step_3d", "oracle_vert_diffuse_3d",
                   "oracle_tr_bounds"):
            getattr(L, fn).argtypes = [G, W, P]
            getattr(L, fn).restype = None
        L.oracle_set_lateral_BC_tr.argtypes = [G, W]
        L.oracle_advect_diffuse.argtypes = [G, W, P]
        L.oracle_advect_diffuse.restype = C.c_int
        for fn in ("oracle_advect_diffuse_mt", "oracle_vert_diffuse_3d_mt",
                   "oracle_tracer_step_3d_mt"):
            getattr(L, fn).argtypes = [G, W, P, C.c_int]
            getattr(L, fn).restype = C.c_int
        for fn in ("oracle_reset_dz", "oracle_calc_age", "oracle_do_ts_relax",
                   "oracle_auxiliary_routines"):
            getattr(L, fn).argtypes = [G, W, P]
            getattr(L, fn).restype = None
        for fn in ("oracle_set_flux_3d", "oracle_vel_w_update"):
            getattr(L, fn).argtypes = [G, W, P]
            getattr(L, fn).restype = None
        L.oracle_set_dz_at_u.argtypes = [G, W, C.c_int]
        L.oracle_set_dz_at_u.restype = None
        L.oracle_transport_step_mt.argtypes = [G, W, P, C.c_int]
        L.oracle_transport_step_mt.restype = C.c_int
        L.oracle_check_tmass.argtypes = [G, W, P, C.c_int]
        L.oracle_check_tmass.restype = C.c_double
        L.oracle_max_threads.restype = C.c_int
        L.oracle_ffsl_maps.argtypes = [C.POINTER(C.POINTER(C.c_int)),
                                       C.POINTER(C.POINTER(C.c_double))]
        L.oracle_last_error.restype = C.c_int
        _lib = L
    return _lib


def set_dz(gp, wp, pp):
    lib().oracle_set_dz(gp, wp, pp)


def step(case, nthreads=1):
    return lib().oracle_tracer_step_3d_mt(case.gp, case.wp, case.pp, nthreads)


def advect_diffuse(case, nthreads=1, lateral=True):
    """advect_diffuse as tracer_step_3d calls it: tracer_step fills the
    lateral ghosts first (tracers.c:153-156)."""
    if lateral:
        lib().oracle_set_lateral_BC_tr(case.gp, case.wp)
    return lib().oracle_advect_diffuse_mt(case.gp, case.wp, case.pp, nthreads)


def vert_diffuse(case, nthreads=1):
    return lib().oracle_vert_diffuse_3d_mt(case.gp, case.wp, case.pp, nthreads)


def tmass(case, n):
    return lib().oracle_check_tmass(case.gp, case.wp, case.pp, n)


def ffsl_maps(case):
    """FFSL trajectory arrays of the last sub-step: (cl, crf), each a list
    of six numpy copies [x face, y face, z face, x centre, y centre,
    z centre] over the sparse array size."""
    cl = (C.POINTER(C.c_int) * 6)()
    crf = (C.POINTER(C.c_double) * 6)()
    lib().oracle_ffsl_maps(cl, crf)
    n = case.sgsiz
    return ([np.ctypeslib.as_array(cl[i], shape=(n,)).copy() for i in range(6)],
            [np.ctypeslib.as_array(crf[i], shape=(n,)).copy() for i in range(6)])


def set_flux_3d(case):
    """set_flux_3d (model/hd/momentum/vel3d.c:4041-4191, VEL3D mode)."""
    lib().oracle_set_flux_3d(case.gp, case.wp, case.pp)


def vel_w_update(case):
    """vel_w_update's continuity loop (vel3d.c:4680-4711); needs set_dz."""
    lib().oracle_vel_w_update(case.gp, case.wp, case.pp)


def set_dz_at_u(case, direction):
    """set_dz_at_u1 (direction 1) / set_dz_at_u2 (2), vel3d.c:760-852."""
    lib().oracle_set_dz_at_u(case.gp, case.wp, direction)


def transport_step(case, nthreads=1):
    """One offline transport step (transport_step + tracer_step +
    transport_post) from the stored flow in the window."""
    return lib().oracle_transport_step_mt(case.gp, case.wp, case.pp, nthreads)


def auxiliary_routines(case):
    """auxiliary_routines' state changes (reset_dz, calc_age, tracer_decay,
    do_ts_relax); returns the oracle's error code (5: a DE_TR2 decay)."""
    L = lib()
    L.oracle_auxiliary_routines(case.gp, case.wp, case.pp)
    return L.oracle_last_error()
