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
    best, trials = autotune.tune_context(tc, blocks=[16], strips=[64, 16, 32, 256],
                                         bins=[1, 3, 8])
    tc.close()
    assert len(trials) == 6 and best in trials
    assert all(t[3] > 0 for t in trials)
    out = []
    snap = H.snapshot(case)
    for shape in (None, best[:3], (16, 16, 1), (16, 256, 2), (16, 32, 5), (16, 64, 64)):
        H.restore(case, snap)
        c = T.Context(case, T.BATCH_F32)
        c.upload()
        if shape:
            c.tune(shape[0], shape[1], shape[2])
            tx, sw, nb, gf = autotune.shape(c)
            assert (tx, sw, nb) == tuple(shape) and gf % 8 == 0
        c.step(6)
        c.download()
        out.append(case.tracers())
        c.close()
    for o in out[1:]:
        assert np.array_equal(out[0], o)


def test_bad_block_width_rejected():
    case = configs.small(seed=6, nxi=20, nyi=16, nz=5, ntr=2)
    L = _lib.lib()
    L.set_dz(case.gp, case.wp, case.pp)
    L.emstr_forget_window(case.gp)
    c = T.Context(case, T.BATCH_F32)
    with pytest.raises(T.EmstrError, match="block width"):
        c.tune(12, 0, 0)
    c.close()
