"""emstr_tracer_mass_noalloc (the entry point that allocates nothing: its
partials in the library's fixed device scratch) against
emstr_tracer_mass on the same contexts, bit for bit, in every form; two
calls in a row, the second on a smaller window, then the larger one
again (the scratch keeps the last call's contents)."""
import ctypes as C
import sys

import numpy as np
import pytest

from ems_amd import _lib, configs, structs as S
from ems_amd import tracer as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(ctx):
    m = np.zeros(ctx.case.ntr)
    r = _lib.lib().emstr_tracer_mass_noalloc(ctx.ptr, m.ctypes.data_as(C.POINTER(C.c_double)))
    assert r == 0, _lib.last_error()
    return m


def _ctx(seed, nxi, nyi, nz, ntr, form):
    case = configs.small(seed=seed, nxi=nxi, nyi=nyi, nz=nz, ntr=ntr, trasc=S.VANLEER)
    case.prepare(O.set_dz)
    ctx = T.Context(case, form)
    ctx.upload()
    return case, ctx


@pytest.mark.parametrize("form", [T.PLAIN_F64, T.PREC_F32, T.BATCH_F32])
def test_noalloc_mass_equals_mass(form):
    say("case no-alloc mass, form", form)
    big = _ctx(31, 40, 36, 8, 5, form)
    small = _ctx(32, 12, 10, 4, 3, form)
    for case, ctx in (big, small, big):
        ctx.step(2)
        ctx.sync()
        want = ctx.mass()
        got = _noalloc(ctx)
        assert np.array_equal(got, want), (got, want)
        assert np.all(np.isfinite(got)) and np.all(got != 0.0)
    for case, ctx in (big, small):
        ctx.close()
