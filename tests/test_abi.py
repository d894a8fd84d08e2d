"""The C-ABI library loads and exports every function include/*.h declares."""
import ctypes as C
import os
import re

from ems_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(headers=("emstr.h", "emstr_grid.h")):
    names = set()
    for h in headers:
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", src, re.M):
            if m.group(0).lstrip().startswith(("typedef", "#")):
                continue
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("tracer_step_3d", "advect_diffuse", "vert_diffuse_3d", "set_dz",
              "emstr_create", "emstr_step", "emstr_grid_build"):
        assert n in names


def test_library_exports_all_declared_symbols():
    L = C.CDLL(_lib.LIBPATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing


def test_prototype_table_matches_header():
    names = declared_functions()
    table = {p[0] for p in _lib.PROTOS}
    assert names == table, (names ^ table)


US_HEADERS = ("emstr_us.h", "emstr_mesh.h")


def test_compas_library_exports_all_declared_symbols():
    from ems_amd import mesh
    names = declared_functions(US_HEADERS)
    for n in ("tracer_step_3d", "advect_diffuse", "vert_diffuse_3d", "set_dz",
              "emstr_mesh_build", "emstr_us_forget_window"):
        assert n in names
    L = C.CDLL(mesh.LIBPATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
