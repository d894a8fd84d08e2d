"""ctypes wrapper of the COMPAS CPU oracle (oracle/liboracle_us.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle_us.so")
_lib = None


def build():
    r = subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle_us.so"],
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIBPATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        from ems_amd import structs_us as U
        L = C.CDLL(LIBPATH)
        G, W, P = C.POINTER(U.Geometry), C.POINTER(U.Window), C.POINTER(U.WinPriv)
        for fn in ("oracle_us_set_dz", "oracle_us_tracer_step_3d",
                   "oracle_us_vert_diffuse_3d", "oracle_us_tr_bounds"):
            getattr(L, fn).argtypes = [G, W, P]
            getattr(L, fn).restype = None
        L.oracle_us_set_lateral_BC_tr.argtypes = [G, W]
        L.oracle_us_advect_diffuse.argtypes = [G, W, P]
        L.oracle_us_advect_diffuse.restype = C.c_int
        for fn in ("oracle_us_advect_diffuse_mt", "oracle_us_vert_diffuse_3d_mt",
                   "oracle_us_tracer_step_3d_mt"):
            getattr(L, fn).argtypes = [G, W, P, C.c_int]
            getattr(L, fn).restype = C.c_int
        L.oracle_us_tmass.argtypes = [G, W, P, C.c_int]
        L.oracle_us_tmass.restype = C.c_double
        L.oracle_us_last_error.restype = C.c_int
        _PD = C.POINTER(C.c_double)
        L.oracle_us_ffsl_posc.argtypes = [G, C.c_int, C.c_int, C.c_double, C.c_double,
                                          C.c_double, _PD]
        L.oracle_us_ffsl_posc.restype = C.c_int
        L.oracle_us_ffsl_limit.argtypes = [G, P, _PD, _PD, _PD]
        L.oracle_us_ffsl_limit.restype = None
        _PI = C.POINTER(C.c_int)
        L.oracle_us_ffsl_maps.argtypes = [G, P, _PI, _PI, _PI]
        L.oracle_us_ffsl_maps.restype = None
        L.oracle_us_set_vlsq.argtypes = [C.c_void_p]
        L.oracle_us_set_vlsq.restype = None
        L.oracle_us_forget_fit.restype = None
        _lib = L
    return _lib


def _vlsq(case):
    """Hand an FFSL window's streamline velocity weights to the oracle:
    the table the library builds from the mesh (window geometry, as the
    least-squares weights of the tracers are)."""
    from ems_amd import structs_us as U
    if not getattr(case, "_oracle_seen", False):
        lib().oracle_us_forget_fit()
        case._oracle_seen = True
    if case.wc.trasc == U.FFSL and hasattr(case, "vlsq"):
        lib().oracle_us_set_vlsq(C.addressof(case.vlsq()))


def set_dz(gp, wp, pp):
    lib().oracle_us_set_dz(gp, wp, pp)


def step(case, nthreads=1):
    _vlsq(case)
    return lib().oracle_us_tracer_step_3d_mt(case.gp, case.wp, case.pp, nthreads)


def advect_diffuse(case, nthreads=1, lateral=True):
    """advect_diffuse as tracer_step_3d calls it, after tracer_step has
    filled the lateral ghosts (tracers.c:185-200)."""
    if lateral:
        lib().oracle_us_set_lateral_BC_tr(case.gp, case.wp)
    _vlsq(case)
    return lib().oracle_us_advect_diffuse_mt(case.gp, case.wp, case.pp, nthreads)


def vert_diffuse(case, nthreads=1):
    return lib().oracle_us_vert_diffuse_3d_mt(case.gp, case.wp, case.pp, nthreads)


def tmass(case, n):
    return lib().oracle_us_tmass(case.gp, case.wp, case.pp, n)


def ffsl_posc(case, e, ci, u, v, dt):
    """get_posc of edge e from its upstream cell ci: (signed cell, x, y,
    distance)."""
    out = np.zeros(3)
    r = lib().oracle_us_ffsl_posc(case.gp, int(e), int(ci), float(u), float(v), float(dt),
                                  out.ctypes.data_as(C.POINTER(C.c_double)))
    return r, out[0], out[1], out[2]


def ffsl_limit(case, tr):
    """get_linear_limit of the sparse array tr: weights [szc, 3], beta [szc]."""
    t = np.ascontiguousarray(tr, dtype=np.float64)
    w = np.zeros(3 * case.szc)
    b = np.zeros(case.szc)
    P = C.POINTER(C.c_double)
    lib().oracle_us_ffsl_limit(case.gp, case.pp, t.ctypes.data_as(P), w.ctypes.data_as(P),
                               b.ctypes.data_as(P))
    return w.reshape(case.szc, 3), b


def ffsl_maps(case):
    """The FFSL maps of the oracle's last sub-step (single thread):
    (cl [sze], clzf [szc], clzc [szc]), 0 where not set."""
    cl = np.zeros(case.sze, dtype=np.int32)
    cf = np.zeros(case.szc, dtype=np.int32)
    cc = np.zeros(case.szc, dtype=np.int32)
    P = C.POINTER(C.c_int)
    lib().oracle_us_ffsl_maps(case.gp, case.pp, cl.ctypes.data_as(P), cf.ctypes.data_as(P),
                              cc.ctypes.data_as(P))
    return cl, cf, cc
