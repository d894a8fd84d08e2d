"""The batched form's autotuner (ems_amd/autotune.py): it times candidate
launch shapes and applies the fastest; results must not depend on the
shape, so a tuned context, a default one and a deliberately odd shape
agree bit for bit after several steps."""
import sys

import numpy as np
import pytest

from ems_amd import _lib, autotune, configs
from ems_amd import tracer as T
from tests import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_tuned_shape_keeps_results():
    say("autotune on a 60x50x12 window, 8 tracers")
    case = configs.small(seed=5, nxi=60, nyi=50, nz=12, ntr=8)
    L = _lib.lib()
    L.set_dz(case.gp, case.wp, case.pp)
    L.emstr_forget_window(case.gp)
    p = T.Context(case, T.PLAIN_F64)
    p.upload_fields()
    p.vel_w()
    p.close()
    tc = T.Context(case, T.BATCH_F32)
    tc.upload()
    best, trials = autotune.tune_context(tc, tiles=[(32, 32), (8, 8), (16, 32)],
                                         scales=(1.0, 0.5))
    full = autotune.shape(tc)[3]
    tc.close()
    assert len(trials) == 4 and best in trials
    assert all(t[3] > 0 for t in trials)
    out = []
    snap = H.snapshot(case)
    for shape in (None, best[:3], (8, 8, 8)):
        H.restore(case, snap)
        c = T.Context(case, T.BATCH_F32)
        c.upload()
        if shape:
            c.tune(*shape)
            tx, ty, g, gf = autotune.shape(c)
            assert (tx, ty, g) == tuple(shape) and gf == full
        c.step(6)
        c.download()
        out.append(case.tracers())
        c.close()
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], out[2])
