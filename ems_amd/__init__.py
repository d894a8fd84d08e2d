"""MI355X-native EMS tracer transport step (SHOC structured path).

The product is libemstr.so: HIP kernels for gfx950 behind a C-ABI that
keeps the reference's entry points (tracer_step_3d, advect_diffuse,
vert_diffuse_3d, set_dz) and structures (geometry_t, window_t,
win_priv_t); see include/emstr.h.  This package loads it with ctypes and
provides seeded synthetic windows (configs) and a device-context wrapper
(tracer.Context).
"""
from ._lib import lib, last_error  # noqa: F401
