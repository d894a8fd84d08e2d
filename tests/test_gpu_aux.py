"""GPU auxiliary_routines (reset_dz, calc_age, tracer_decay, do_ts_relax;
model/hd/tracers/tracers.c:4296-4490) through the C-ABI against the
oracle, bit for bit: the drop-in auxiliary_routines alone, and the
drop-in tracer_step_3d (transport, then auxiliary_routines) over steps
with a moving surface."""
import ctypes as C
import sys

import numpy as np
import pytest

from ems_amd import _lib, structs as S
from ems_amd import tracer as T
from oracle import oracle as O
from tests.test_oracle_aux import aux_case, set_ptr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def with_processes(seed):
    """A moving-surface window with every built process on: decay of
    tracer 1 (constant) and 2 (timescale from tracer 3, hours),
    temperature (0) and salinity (4) relaxation with a no-gradient field,
    time constants for temperature, and an age tracer (5) in a mask."""
    case, rng = aux_case(seed)
    wd, wc = case.wd, case.wc
    ti = wc.trinfo_3d
    tdec = S.view(wc.tdec, case.ntr, np.int32)
    dectr = S.view(wc.dectr, case.ntr, np.float64)
    wc.ntdec = 2
    tdec[:2] = [1, 2]
    dectr[1] = 6.0 * 3600.0
    dectr[2] = 3.0
    ti[2].flag = S.DE_TR3
    ti[3].units = b"hours"
    wd.tno, wd.sno = 0, 4
    set_ptr(case, wd, "rtemp", rng.uniform(1.0, 5.0, case.sgsiz))
    set_ptr(case, wd, "rsalt", rng.uniform(1.0, 5.0, case.sgsiz))
    set_ptr(case, wd, "temp_tc", rng.uniform(0.5, 3.0, case.sgsiz))
    ti[0].relax_rate = 3e-5
    ti[0].tctype = S.RLX_ADPT
    ti[0].flag = S.RLX_GRD
    ti[0].relax_dum = -6.0
    ti[4].relax_rate = 1e-5
    ti[4].flag = S.RLX_GRD
    ti[4].relax_dum = -3.0
    wd.agetr = wd.tr_wc[5]
    set_ptr(case, wc, "agemsk", (rng.uniform(size=case.sgsiz) < 0.6).astype(np.int16),
            C.c_short)
    return case


def state(case):
    wc = case.wc
    s2 = S.view(wc.s2, case.sgsiz, np.int32)[1:wc.vc2 + 1].copy()
    c1 = np.frombuffer(C.string_at(wc.c1, case.sgsiz), np.uint8).copy()
    return dict(tr=case.tracers(), dz=case.a3(wc.dz)[case.wet].copy(), s2=s2, c1=c1,
                lists=(wc.vc2, wc.vcs2, wc.vca2, wc.vci2),
                rtemp=case.a3(case.wd.rtemp)[case.wet].copy(),
                rsalt=case.a3(case.wd.rsalt)[case.wet].copy())


def check_equal(a, b):
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), (k, np.abs(a[k] - b[k]).max())
        else:
            assert a[k] == b[k], k


@pytest.mark.parametrize("seed", [21, 22])
def test_auxiliary_routines_bit_exact(seed):
    say("auxiliary_routines drop-in vs oracle", seed)
    a, b = with_processes(seed), with_processes(seed)
    L = _lib.lib()
    for _ in range(2):
        L.auxiliary_routines(a.gp, a.wp, a.pp)
        msg, cell = _lib.last_error()
        assert msg == "", (msg, cell)
        assert O.auxiliary_routines(b) == 0
    L.emstr_forget_window(a.gp)
    check_equal(state(a), state(b))


def test_tracer_step_3d_with_auxiliary_bit_exact():
    """Three steps of the time loop (hd_step: the surface moves, set_dz,
    then tracer_step_3d, i.e. the transport and auxiliary_routines): the
    drop-in set_dz and tracer_step_3d against the oracle's, bit for bit.
    (set_dz rewrites the bottom cells i1 that reset_dz reuses for its own
    list; the GPU reads the bed from the grid.)"""
    say("tracer_step_3d with auxiliary processes vs oracle")
    a, b = with_processes(23), with_processes(23)
    L = _lib.lib()
    rng = np.random.default_rng(6)
    w2 = S.view(a.g.w2_t, a.sgsizS, np.int32)[1:a.g.v2_t + 1]
    dl = a.layers[1] - a.layers[0]
    for s in range(3):
        if s:
            step = rng.uniform(-0.4, 0.4, len(w2)) * dl
            for c in (a, b):
                eta, old = c.a2(c.wd.eta), c.a2(c.wc.oldeta)
                old[w2] = eta[w2]
                eta[w2] += step
            L.set_dz(a.gp, a.wp, a.pp)
            O.set_dz(b.gp, b.wp, b.pp)
        L.tracer_step_3d(a.gp, a.wp, a.pp)
        msg, cell = _lib.last_error()
        assert msg == "", (msg, cell)
        assert O.step(b) == 0
    L.emstr_forget_window(a.gp)
    check_equal(state(a), state(b))


def test_fp32_forms_refuse_tracer_processes():
    say("auxiliary processes refused by the FP32 forms")
    case = with_processes(24)
    L = _lib.lib()
    ctx = T.Context(case, T.BATCH_F32)
    ctx.upload()
    assert L.emstr_auxiliary(ctx.ptr) == 2
    assert "plain form only" in _lib.last_error()[0]
    case.wc.ntdec = 0
    case.wd.tno = case.wd.sno = -1
    case.wd.agetr = None
    assert L.emstr_auxiliary(ctx.ptr) == 0
    ctx.close()
