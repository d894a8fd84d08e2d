"""The drop-in boundary without a GPU: a C caller compiled against the
public headers links the libraries and reaches their entry points; the
paths the libraries do not cover (open boundaries, sigma coordinates,
2D mode, FFSL fills) are rejected with an error and never exit; the
arguments of tracer_step are checked.  None of these calls reaches the
device."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from ems_amd import _lib, configs, mesh as MS, structs as S, structs_us as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "ems_amd")

SHOC_CALLER = r"""
#include <stdio.h>
#include <string.h>
#include "emstr.h"

/* every entry point of emstr.h, as the reference's callers bind them */
typedef void (*step_fn)(geometry_t *, window_t *, win_priv_t *);
int main(void) {
  void *eps[] = {(void *)tracer_step, (void *)tracer_step_3d, (void *)advect_diffuse,
                 (void *)vert_diffuse_3d, (void *)set_dz, (void *)emstr_create,
                 (void *)emstr_steps, (void *)emstr_tracer_mass,
                 (void *)emstr_tracer_mass_noalloc, (void *)emstr_forget_window};
  for (unsigned i = 0; i < sizeof(eps) / sizeof(eps[0]); i++)
    if (!eps[i]) return 10 + (int)i;
  master_t m;
  memset(&m, 0, sizeof(m));
  m.mode2d = 1;
  geometry_t g;
  window_t wd;
  win_priv_t wc;
  memset(&g, 0, sizeof(g));
  memset(&wd, 0, sizeof(wd));
  memset(&wc, 0, sizeof(wc));
  geometry_t *gw[2] = {NULL, &g};
  window_t *dw[2] = {NULL, &wd};
  win_priv_t *cw[2] = {NULL, &wc};
  tracer_step(&m, gw, dw, cw, 1);
  if (m.crf != RS_RESTART || !strstr(emstr_last_error(), "2D mode")) return 2;
  /* an open-boundary window is refused before anything reaches the GPU */
  g.nobc = 2;
  if (emstr_create(&g, &wd, &wc, EMSTR_PLAIN_F64) != NULL) return 3;
  if (!strstr(emstr_last_error(), "open boundaries")) return 4;
  ((step_fn)tracer_step_3d)(&g, &wd, &wc);
  if (!strstr(emstr_last_error(), "open boundaries")) return 5;
  printf("caller ok: %s\n", emstr_last_error());
  return 0;
}
"""

COMPAS_CALLER = r"""
#include <stdio.h>
#include <string.h>
#include "emstr_us.h"

int main(void) {
  void *eps[] = {(void *)tracer_step, (void *)tracer_step_3d, (void *)advect_diffuse,
                 (void *)vert_diffuse_3d, (void *)set_dz, (void *)emstr_us_create,
                 (void *)emstr_us_steps, (void *)emstr_us_forget_window};
  for (unsigned i = 0; i < sizeof(eps) / sizeof(eps[0]); i++)
    if (!eps[i]) return 10 + (int)i;
  master_t m;
  memset(&m, 0, sizeof(m));
  tracer_step(&m, NULL, NULL, NULL, 1);
  if (m.crf != RS_RESTART || !strstr(emstr_us_last_error(), "null structure")) return 2;
  geometry_t g;
  window_t wd;
  win_priv_t wc;
  memset(&g, 0, sizeof(g));
  memset(&wd, 0, sizeof(wd));
  memset(&wc, 0, sizeof(wc));
  wc.sigma = 1;
  if (emstr_us_create(&g, &wd, &wc) != NULL) return 3;
  if (!strstr(emstr_us_last_error(), "sigma")) return 4;
  printf("caller ok: %s\n", emstr_us_last_error());
  return 0;
}
"""


def _build_and_run(tmp_path, src, lib):
    c = tmp_path / "caller.c"
    c.write_text(src)
    exe = tmp_path / "caller"
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           str(c), "-o", str(exe), "-L", LIBDIR, "-l" + lib, "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "caller ok" in r.stdout


def test_c_caller_links_shoc_library(tmp_path):
    _build_and_run(tmp_path, SHOC_CALLER, "emstr")


def test_c_caller_links_compas_library(tmp_path):
    _build_and_run(tmp_path, COMPAS_CALLER, "emstr_us")


@pytest.mark.parametrize("what", ["nobc", "sigma"])
def test_shoc_out_of_scope_windows_rejected(what):
    case = configs.small(seed=1, nxi=8, nyi=6, nz=3, ntr=1)
    L = _lib.lib()
    if what == "nobc":
        # open boundaries: the plain form only (and a boundary structure)
        case.open_boundary(S.L_EDGE)
        want = "open boundaries"
        assert not L.emstr_create(case.gp, case.wp, case.pp, 1)
        assert want in _lib.last_error()[0]
        case.g.nobc, keep = 1, case.g.open
        case.g.open = None
    else:
        case.wc.sigma = 1
        want = "sigma"
    assert not L.emstr_create(case.gp, case.wp, case.pp, 0)
    msg, _ = _lib.last_error()
    assert want in msg
    # the drop-in entry points record the same error and return
    snap = case.tracers()
    L.tracer_step_3d(case.gp, case.wp, case.pp)
    assert want in _lib.last_error()[0]
    assert L.advect_diffuse(case.gp, case.wp, case.pp) != 0
    assert np.array_equal(case.tracers(), snap)
    if what == "nobc":
        case.g.open = keep


def test_shoc_tracer_step_arguments():
    L = _lib.lib()
    m = S.Master()
    L.tracer_step(C.byref(m), None, None, None, 1)
    assert m.crf == S.RS_RESTART and "null structure" in _lib.last_error()[0]
    case = configs.small(seed=1, nxi=8, nyi=6, nz=3, ntr=1)
    G = (C.POINTER(S.Geometry) * 2)(None, case.gp)
    W = (C.POINTER(S.Window) * 2)(None, case.wp)
    P = (C.POINTER(S.WinPriv) * 2)(None, case.pp)
    m = S.Master(mode2d=1)
    L.tracer_step(C.byref(m), G, W, P, 1)
    assert m.crf == S.RS_RESTART and "2D mode" in _lib.last_error()[0]
    case.g.nobc = 1
    m = S.Master()
    L.tracer_step(C.byref(m), G, W, P, 1)
    assert m.crf == S.RS_RESTART and "open boundaries" in _lib.last_error()[0]


@pytest.mark.parametrize("what", ["nobc", "sigma", "fill", "rk4"])
def test_compas_out_of_scope_windows_rejected(what):
    hc = MS.hex_basin(nx=8, ny=6, nz=3, ntr=1, seed=1,
                      trasc=U.FFSL if what == "fill" else U.VANLEER)
    L = MS.lib()
    if what == "nobc":
        hc.g.nobc = 1
        want = "open boundaries"
    elif what == "sigma":
        hc.wc.sigma = 1
        want = "sigma"
    elif what == "fill":
        hc.wc.fillf = U.MONOTONIC
        want = "MONOTONIC without CLIP"
    else:
        hc.wc.rkstage = 4
        want = "4 stages"
    assert not L.emstr_us_create(hc.gp, hc.wp, hc.pp)
    assert want in MS.last_error()[0]
    L.tracer_step_3d(hc.gp, hc.wp, hc.pp)
    assert want in MS.last_error()[0]
