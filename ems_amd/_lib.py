"""Loads libemstr.so (the C-ABI library built from ems_amd/csrc) via ctypes.

The library holds the HIP kernels and their C host layer.  There is no
fallback: if the library is missing, loading raises.  EMSTR_LIB, if set,
names another build of the library to load instead of the in-tree one.
"""
import ctypes as C
import os

from . import structs as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("EMSTR_LIB") or os.path.join(HERE, "libemstr.so")

_lib = None

# Exported symbols and their prototypes: (name, restype, argtypes).
_G = C.POINTER(S.Geometry)
_W = C.POINTER(S.Window)
_P = C.POINTER(S.WinPriv)
_CTX = C.c_void_p
PROTOS = [
    # drop-in entry points (include/emstr.h)
    ("tracer_step", None, [C.POINTER(S.Master), C.POINTER(_G), C.POINTER(_W), C.POINTER(_P),
                           C.c_int]),
    ("tracer_step_3d", None, [_G, _W, _P]),
    ("advect_diffuse", C.c_int, [_G, _W, _P]),
    ("vert_diffuse_3d", None, [_G, _W, _P]),
    ("set_dz", None, [_G, _W, _P]),
    ("auxiliary_routines", None, [_G, _W, _P]),
    # device-resident interface
    ("emstr_create", _CTX, [_G, _W, _P, C.c_int]),
    ("emstr_destroy", None, [_CTX]),
    ("emstr_upload_fields", C.c_int, [_CTX]),
    ("emstr_upload_tracers", C.c_int, [_CTX]),
    ("emstr_download_tracers", C.c_int, [_CTX]),
    ("emstr_step", C.c_int, [_CTX]),
    ("emstr_steps", C.c_int, [_CTX, C.c_int]),
    ("emstr_tune", C.c_int, [_CTX, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("emstr_launch_shape", C.c_int, [_CTX] + [C.POINTER(C.c_int)] * 4),
    ("emstr_advect_diffuse", C.c_int, [_CTX]),
    ("emstr_vert_diffuse", C.c_int, [_CTX]),
    ("emstr_tracer_mass", C.c_int, [_CTX, C.POINTER(C.c_double)]),
    ("emstr_tracer_mass_noalloc", C.c_int, [_CTX, C.POINTER(C.c_double)]),
    ("emstr_auxiliary", C.c_int, [_CTX]),
    ("emstr_device_set_dz", C.c_int, [_CTX]),
    ("emstr_device_vel_w", C.c_int, [_CTX]),
    ("emstr_device_set_flux_3d", C.c_int, [_CTX]),
    ("emstr_transport_step", C.c_int, [_CTX]),
    ("emstr_transport_download", C.c_int, [_CTX]),
    ("emstr_profile", C.c_int, [_CTX, C.c_int]),
    ("emstr_kernel_times", C.c_int, [_CTX, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    ("emstr_sync", C.c_int, [_CTX]),
    ("emstr_stream", C.c_void_p, [_CTX]),
    ("emstr_ensemble_create", C.c_void_p, [C.POINTER(_CTX), C.c_int]),
    ("emstr_ensemble_steps", C.c_int, [C.c_void_p, C.c_int]),
    ("emstr_ensemble_sync", C.c_int, [C.c_void_p]),
    ("emstr_ensemble_stream", C.c_void_p, [C.c_void_p]),
    ("emstr_ensemble_destroy", None, [C.c_void_p]),
    ("emstr_forget_window", None, [_G]),
    ("emstr_last_error", C.c_char_p, []),
    ("emstr_last_error_cell", C.c_int, []),
    # grid builder (include/emstr_grid.h)
    ("emstr_grid_build", C.POINTER(S.Grid),
     [C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 9 + [C.c_int]),
    ("emstr_grid_free", None, [C.POINTER(S.Grid)]),
    ("emstr_grid_open", C.c_int, [C.POINTER(S.Grid), C.c_int]),
    ("emstr_vel_w_update", None, [_G, _W, _P]),
]


def lib():
    """Return the loaded library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                "libemstr.so is not built (%s); run __graft_entry__.build()" % LIBPATH)
        L = C.CDLL(LIBPATH)
        for name, res, args in PROTOS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    L = lib()
    return L.emstr_last_error().decode(errors="replace"), L.emstr_last_error_cell()
