#!/usr/bin/env python3
"""Tracer cell-updates/s of the EMS tracer step on one MI355X.

Workload: the 400x600x40 curvilinear basin with 40 tracers (van Leer,
dt = 90 s, seeded; 25 % land, 2-300 m), built by ems_amd.configs.c3().
One step = advect_diffuse + vert_diffuse_3d over all tracers with fields
and tracers resident in HBM.  The FP32 precision form is the headline
value; the FP64 plain form runs on the same inputs for the ratio and as
the accuracy reference (it agrees bit for bit with the CPU oracle).

Prints one JSON line (rank 0, N = 1).  --gpus N with N > 1 exits: this
build runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("tracer cell-updates/s on 1 GPU, FP32-batched vs FP64 plain, at rel err "
          "≤1e-5")


class HipEvents:
    """hipEvent timing on a given stream (the library's context stream)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        t = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(t), a, b) == 0
        return float(t.value)


def timed_steps(ctx, ev, steps):
    a, b = ev.event(), ev.event()
    ctx.sync()
    ev.record(a, ctx.stream)
    ctx.step(steps)
    ev.record(b, ctx.stream)
    ms = ev.elapsed_ms(a, b)
    ctx.sync()
    return ms


def kernel_times(ctx, steps):
    from ems_amd import _lib
    L = _lib.lib()
    L.emstr_profile(ctx.ptr, 1)
    ctx.step(steps)
    ctx.sync()
    ms = (C.c_double * 5)()
    cnt = (C.c_longlong * 5)()
    L.emstr_kernel_times(ctx.ptr, ms, cnt)
    L.emstr_profile(ctx.ptr, 0)
    return list(ms), list(cnt)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--ntr", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the tracer step on one GPU only; "
              "multi-GPU sharding is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from ems_amd import _lib, configs
    from ems_amd import tracer as T

    L = _lib.lib()
    t0 = time.time()
    case = configs.c3(seed=args.seed, ntr=args.ntr)
    print("bench: c3 built in %.1f s: %d wet cells x %d tracers"
          % (time.time() - t0, case.nwet, case.ntr), file=sys.stderr, flush=True)
    # set_dz (device) fills the host lists; w from continuity on the device
    L.set_dz(case.gp, case.wp, case.pp)
    msg, cell = _lib.last_error()
    if msg:
        raise RuntimeError("set_dz: %s" % msg)
    L.emstr_forget_window(case.gp)
    ctx64 = T.Context(case, T.PLAIN_F64)
    ctx64.upload_fields()
    ctx64.vel_w()
    ctx64.upload()
    ctx32 = T.Context(case, T.PREC_F32)
    ctx32.upload()
    m64_0, m32_0 = ctx64.mass(), ctx32.mass()
    ev = HipEvents()
    nwet, ntr = case.nwet, case.ntr
    K, W = args.steps, args.warmup

    ctx64.step(W)
    ms64 = timed_steps(ctx64, ev, K)
    ctx32.step(W)
    ms32 = timed_steps(ctx32, ev, K)
    print("bench: fp64 %.4f ms/step, fp32 %.4f ms/step" % (ms64 / K, ms32 / K),
          file=sys.stderr, flush=True)
    PK = 20
    kt32, kc32 = kernel_times(ctx32, PK)
    kt64, kc64 = kernel_times(ctx64, PK)
    nsteps = W + K + PK

    # accuracy of the FP32 form against the FP64 plain form, same steps
    m64_1, m32_1 = ctx64.mass(), ctx32.mass()
    ctx64.download()
    ref = case.tracers()
    ctx32.download()
    f32 = case.tracers()
    span = np.maximum(np.abs(ref).max(axis=1), 1e-30)
    err = (np.abs(f32 - ref).max(axis=1) / span)
    rms = np.sqrt(((f32 - ref) ** 2).mean(axis=1)) / span
    drift32 = np.abs(m32_1 - m32_0) / np.abs(m32_0)
    drift64 = np.abs(m64_1 - m64_0) / np.abs(m64_0)

    cu = float(nwet) * ntr
    v32 = cu * K / (ms32 * 1e-3)
    v64 = cu * K / (ms64 * 1e-3)
    # dominant FP32 kernel and its algorithmic bytes per launch
    names = ["advect", "vdiff", "prep", "other", "fused"]
    dom = int(np.argmax(kt32))
    alg = {0: nwet * ntr * 8 + nwet * 7 * 4,
           1: nwet * ntr * 8 + nwet * 3 * 4,
           2: nwet * (7 * 8 + 10 * 4),
           4: nwet * ntr * 8 + nwet * 10 * 4}.get(dom, nwet * ntr * 8)
    avg_ms = kt32[dom] / max(kc32[dom], 1)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    # step-level algorithmic bytes (N3*ntr*2*s + N3*8*F, F = 7)
    step_bytes32 = nwet * ntr * 8 + nwet * 56
    step_bytes64 = nwet * ntr * 16 + nwet * 56

    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms32 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded c3 generator, seed %d)" % args.seed,
        "config": {"workload": "c3: 400x600x40 curvilinear basin, 25%% land, 2-300 m, "
                               "40 tracers, van Leer, dt=90 s",
                   "n3_wet": nwet, "ntr": ntr, "columns": int(case.g.v2_t)},
        "fp64_plain": {"value": v64, "ms_per_step": ms64 / K,
                       "step_hbm_fraction": step_bytes64 / (ms64 / K * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "ratio_fp32_over_fp64": v32 / v64,
        "step_hbm_fraction": step_bytes32 / (ms32 / K * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "accuracy": {"steps": nsteps, "max_rel_err": float(err.max()),
                     "rms_rel_err": float(rms.max()),
                     "mass_drift_fp32": float(drift32.max()),
                     "mass_drift_fp64": float(drift64.max())},
        "kernels_ms_per_step": {"fp32": {names[i]: kt32[i] / PK for i in range(5) if kc32[i]},
                                "fp64": {names[i]: kt64[i] / PK for i in range(5) if kc64[i]}},
        "roofline": {"bound": "hbm", "kernel": "fp32 " + names[dom],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    ctx64.close()
    ctx32.close()

    if not args.no_cpu:
        # CPU baseline: the oracle (FP64 C restatement) with OpenMP over
        # tracers on the host cores, a bounded number of steps.
        from oracle import oracle as O
        nth = min(16, os.cpu_count() or 1)
        O.lib()
        O.set_dz(case.gp, case.wp, case.pp)
        t = time.time()
        n = 0
        while True:
            O.step(case, nth)
            n += 1
            if time.time() - t > args.cpu_seconds or n >= 50:
                break
        dt = time.time() - t
        out["cpu_baseline"] = {"value": cu * n / dt, "unit": "cell-updates/s",
                               "cores": nth, "kind": "port",
                               "sample": "%d full c3 steps (all %d tracers), OpenMP over "
                                         "tracers" % (n, ntr)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
