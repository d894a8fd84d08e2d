"""Python wrapper of the device-resident C interface (emstr_ctx).

Every call goes to libemstr.so; errors raise EmstrError with the message
and failing cell the library reports.  There is no host fallback.
"""
import ctypes as C

import numpy as np

from . import _lib

PLAIN_F64 = 0
PREC_F32 = 1
BATCH_F32 = 2


class EmstrError(RuntimeError):
    def __init__(self, code, what):
        msg, cell = _lib.last_error()
        super().__init__("%s failed (code %d): %s [cell %d]" % (what, code, msg, cell))
        self.code = code
        self.cell = cell


def _chk(r, what):
    if r:
        raise EmstrError(r, what)


class Context:
    """Device mirror of one window (a configs.Case or raw structures)."""

    def __init__(self, case, form=PLAIN_F64):
        L = _lib.lib()
        self.case = case
        self.form = form
        self.ptr = L.emstr_create(case.gp, case.wp, case.pp, form)
        if not self.ptr:
            raise EmstrError(-1, "emstr_create")

    def set_dz(self):
        _chk(_lib.lib().emstr_device_set_dz(self.ptr), "emstr_device_set_dz")

    def vel_w(self):
        _chk(_lib.lib().emstr_device_vel_w(self.ptr), "emstr_device_vel_w")

    def set_flux_3d(self):
        _chk(_lib.lib().emstr_device_set_flux_3d(self.ptr), "emstr_device_set_flux_3d")

    def transport_step(self):
        _chk(_lib.lib().emstr_transport_step(self.ptr), "emstr_transport_step")

    def transport_download(self):
        _chk(_lib.lib().emstr_transport_download(self.ptr), "emstr_transport_download")

    def upload(self):
        L = _lib.lib()
        _chk(L.emstr_upload_fields(self.ptr), "emstr_upload_fields")
        _chk(L.emstr_upload_tracers(self.ptr), "emstr_upload_tracers")

    def upload_fields(self):
        _chk(_lib.lib().emstr_upload_fields(self.ptr), "emstr_upload_fields")

    def upload_tracers(self):
        _chk(_lib.lib().emstr_upload_tracers(self.ptr), "emstr_upload_tracers")

    def download(self):
        _chk(_lib.lib().emstr_download_tracers(self.ptr), "emstr_download_tracers")

    def step(self, n=1):
        _chk(_lib.lib().emstr_steps(self.ptr, n), "emstr_steps")

    def tune(self, tile_x=0, tile_y=0, grid=0, use_graph=-1):
        _chk(_lib.lib().emstr_tune(self.ptr, tile_x, tile_y, grid, use_graph), "emstr_tune")

    def advect_diffuse(self):
        _chk(_lib.lib().emstr_advect_diffuse(self.ptr), "emstr_advect_diffuse")

    def vert_diffuse(self):
        _chk(_lib.lib().emstr_vert_diffuse(self.ptr), "emstr_vert_diffuse")

    def sync(self):
        _chk(_lib.lib().emstr_sync(self.ptr), "emstr_sync")

    def mass(self):
        m = np.zeros(self.case.ntr)
        _chk(_lib.lib().emstr_tracer_mass(self.ptr, m.ctypes.data_as(C.POINTER(C.c_double))),
             "emstr_tracer_mass")
        return m

    @property
    def stream(self):
        return _lib.lib().emstr_stream(self.ptr)

    def close(self):
        if self.ptr:
            _lib.lib().emstr_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Batched-form contexts stepped together: one hipGraph with a branch
    per context (emstr_ensemble_create); steps must be even."""

    def __init__(self, contexts):
        L = _lib.lib()
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.ptr for c in self.contexts])
        self.ptr = L.emstr_ensemble_create(arr, len(self.contexts))
        if not self.ptr:
            raise RuntimeError("emstr_ensemble_create: %s (cell %d)" % _lib.last_error())
        self.stream = L.emstr_ensemble_stream(self.ptr)

    def step(self, n=2):
        _chk(_lib.lib().emstr_ensemble_steps(self.ptr, n), "emstr_ensemble_steps")

    def sync(self):
        _chk(_lib.lib().emstr_ensemble_sync(self.ptr), "emstr_ensemble_sync")

    def close(self):
        if self.ptr:
            _lib.lib().emstr_ensemble_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
