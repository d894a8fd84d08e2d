#!/usr/bin/env python3
"""Tracer cell-updates/s of the EMS tracer step on one MI355X.

Workload: the 400x600x40 curvilinear basin with 40 tracers (van Leer,
dt = 90 s, seeded; 25 % land, 2-300 m), built by ems_amd.configs.c3().
One step = advect_diffuse + vert_diffuse_3d over all tracers with fields
and tracers resident in HBM, including the per-step coefficient prep.
The batched FP32 form is the headline value; the FP64 plain form runs on
the same inputs for the ratio and as the accuracy reference (it agrees bit
for bit with the CPU oracle); the unfused FP32 precision form is reported
beside them.

Prints one JSON line (rank 0, N = 1).  --gpus N with N > 1 exits: this
build runs on one GPU.

  --forms plain,precision,batched   which forms to run (profiling runs)
  --workload c3-ffsl                the same basin with the structured FFSL
                                    scheme (plain FP64 form; 1 tracer line)
  --workload c4-compas              the COMPAS path on the 350x350 hexagon
                                    mesh, 50 layers, 60 tracers, van Leer
                                    (plain FP64 form)
  --workload c4-compas-ffsl         the same mesh with COMPAS FFSL: the FP32
                                    form beside the FP64 plain form, error
                                    and drift over 500 steps
  --workload c5-ensemble            the 64-window set in the batched FP32
                                    form, one ensemble hipGraph, with the
                                    per-window error and mass drift against
                                    the FP64 plain form
  --workload c2-precision           the 200x300x30 QUICKEST case: the FP32
                                    precision form against the FP64 plain
                                    form, error and drift over 500 steps
  --copy-probe MB                   time a plain device copy of MB MiB and
                                    exit (calibrates the HBM byte counters)
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


def measured_traffic(kernel):
    """HBM bytes per dispatch of a kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy
    calibration recorded there); None if the kernel is not listed."""
    path = os.path.join(ROOT, "profiles", "batched_c3_hbm_traffic.md")
    if kernel is None or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == kernel:
            return int(f[2])
    return None


def copy_probe(mb):
    """A plain device-to-device copy of mb MiB, five times: its
    FETCH_SIZE / WRITE_SIZE per dispatch against the known 2 x mb MiB
    gives the scale factor of those counters."""
    import torch
    n = mb * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    y.copy_(x)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    print(json.dumps({"copy_probe_MiB": mb, "ms": ms,
                      "GBps": 2 * mb * (1 << 20) / (ms * 1e-3) / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--ntr", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the accuracy comparison against the FP64 form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the batched form's default launch shape")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--forms", default="plain,precision,batched")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--copy-probe", type=int, default=0)
    ap.add_argument("--workload", default="c3",
                    choices=["c3", "c3-ffsl", "c4-compas", "c4-compas-ffsl", "c5-ensemble",
                             "c2-precision"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the tracer step on one GPU only; "
              "runs on several GPUs are not part of it", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "c3-ffsl":
        return bench_ffsl(args)
    if args.workload in ("c4-compas", "c4-compas-ffsl"):
        return bench_compas(args)
    if args.workload == "c5-ensemble":
        return bench_ensemble(args)
    if args.workload == "c2-precision":
        return bench_c2(args)
    forms = set(args.forms.split(","))

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
    ev = HipEvents()
    nwet, ntr = case.nwet, case.ntr
    K, W, PK = args.steps, args.warmup, args.profile_steps
    cu = float(nwet) * ntr
    names = ["advect", "vdiff", "prep", "other", "fused"]
    res = {}
    ctx = {}
    tuned = None
    for form, kind in (("plain", T.PLAIN_F64), ("precision", T.PREC_F32),
                       ("batched", T.BATCH_F32)):
        if form not in forms:
            continue
        c = T.Context(case, kind)
        if form == "plain":
            c.upload_fields()
            c.vel_w()
        c.upload()
        if form == "batched" and not args.no_tune:
            # launch shape by measurement, on a throwaway context so that
            # the measured one starts from the same state as the others
            from ems_amd import autotune
            tc = T.Context(case, kind)
            tc.upload()
            best, trials = autotune.tune_context(tc)
            tc.close()
            c.tune(best[0], best[1], best[2])
            tuned = {"block": [best[0], 4], "strip": best[1], "depth_bins": best[2],
                     "ms_per_step": best[3],
                     "trials": [[t[0], t[1], t[2], round(t[3], 4)] for t in trials]}
            print("bench: autotune -> block %d x 4, strip %d, %d depth bins (%.4f ms/step)"
                  % (best[0], best[1], best[2], best[3]), file=sys.stderr, flush=True)
        m0 = c.mass()
        c.step(W)
        ms = timed_steps(c, ev, K)
        kt, kc = kernel_times(c, PK)
        res[form] = {"ms": ms / K, "m0": m0, "kt": kt, "kc": kc}
        ctx[form] = c
        print("bench: %s %.4f ms/step" % (form, ms / K), file=sys.stderr, flush=True)
    nsteps = W + K + PK

    out = {"metric": METRIC, "value": None, "unit": "cell-updates/s", "n_gpus": 1,
           "steps": K, "warmup": W, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded c3 generator, seed %d)" % args.seed,
           "config": {"workload": "c3: 400x600x40 curvilinear basin, 25% land, 2-300 m, "
                                  "40 tracers, van Leer, dt=90 s",
                      "n3_wet": nwet, "ntr": ntr, "columns": int(case.g.v2_t),
                      "form": "batched FP32 (fused advection + column solve, hipGraph)"}}
    if "batched" in res:
        rB = res["batched"]
        out["value"] = cu / (rB["ms"] * 1e-3)
        out["ms_per_step"] = rB["ms"]
        ktB, kcB = rB["kt"], rB["kc"]
        dom = int(np.argmax(ktB))
        # algorithmic bytes per launch of the dominant kernel: the fused
        # step reads and writes every tracer of every wet cell once (a
        # 4-byte head and a 1-byte tail each way) and reads the cell's 12
        # coefficient words; the prep reads the 7 FP64 fields and writes
        # the coefficients
        alg = {4: nwet * (ntr * 10 + 12 * 4),
               2: nwet * (7 * 8 + 12 * 4)}.get(dom, nwet * ntr * 10)
        avg_ms = ktB[dom] / max(kcB[dom], 1)
        achieved = alg / (avg_ms * 1e-3) / 1e9
        kname = {4: "k_step4"}.get(dom)
        out["roofline"] = {"bound": "hbm", "kernel": "batched " + names[dom],
                           "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS,
                           "traffic": measured_traffic(kname),
                           "alg_bytes_per_launch": alg, "avg_launch_ms": avg_ms}
        # per step: the tracer state in and out, the 7 FP64 fields in
        out["step_hbm_fraction"] = (nwet * ntr * 10 + nwet * 56) / (rB["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
        if tuned:
            out["autotune"] = tuned
    kms = {}
    for form, r in res.items():
        kms[form] = {names[i]: r["kt"][i] / PK for i in range(5) if r["kc"][i]}
    out["kernels_ms_per_step"] = kms
    if "plain" in res:
        r64 = res["plain"]
        out["fp64_plain"] = {"value": cu / (r64["ms"] * 1e-3), "ms_per_step": r64["ms"],
                             "step_hbm_fraction": (nwet * ntr * 16 + nwet * 56) /
                             (r64["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if "batched" in res:
            out["ratio_fp32_over_fp64"] = r64["ms"] / res["batched"]["ms"]
    if "precision" in res:
        out["fp32_precision_unfused"] = {"value": cu / (res["precision"]["ms"] * 1e-3),
                                         "ms_per_step": res["precision"]["ms"]}
    if "plain" in res:
        # accuracy of the FP32 forms against the FP64 plain form, every
        # context carried on to the same step count (accuracy_steps)
        c64 = ctx["plain"]
        extra = max(0, args.accuracy_steps - nsteps)
        c64.step(extra)
        m64 = c64.mass()
        c64.download()
        ref = case.tracers()
        for form, key in (("batched", "accuracy"), ("precision", "accuracy_precision")):
            if form not in ctx:
                continue
            cf = ctx[form]
            cf.step(extra)
            mf = cf.mass()
            cf.download()
            out[key] = accuracy(case.tracers(), ref, mf, res[form]["m0"], m64,
                                res["plain"]["m0"], nsteps + extra)
    for c in ctx.values():
        c.close()

    if not args.no_dropin:
        # the drop-in entry point on host buffers (fields uploaded, tracers
        # copied in and out over PCIe every call, plain FP64 form): the
        # PCIe-inclusive rate, reported beside the device-resident value
        L.tracer_step_3d(case.gp, case.wp, case.pp)
        t = time.time()
        nd = 3
        for _ in range(nd):
            L.tracer_step_3d(case.gp, case.wp, case.pp)
        msg, cell = _lib.last_error()
        if msg:
            raise RuntimeError("tracer_step_3d: %s" % msg)
        dms = (time.time() - t) / nd * 1e3
        L.emstr_forget_window(case.gp)
        out["dropin_host_buffers"] = {"ms_per_call": dms, "value": cu / (dms * 1e-3),
                                      "unit": "cell-updates/s", "form": "plain FP64"}

    if not args.no_cpu:
        # CPU baseline: the oracle (FP64 C restatement) with OpenMP over
        # tracers on the host cores, a bounded number of steps.
        from oracle import oracle as O
        O.lib()
        O.set_dz(case.gp, case.wp, case.pp)
        out["cpu_baseline"] = cpu_baselines([case], O.step, args.cpu_seconds, "c3")
    print(json.dumps(out))


def cpu_info():
    """(model name, logical CPUs, CPUs this process may run on)"""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_baselines(cases, step, seconds, what):
    """The oracle (the FP64 C restatement of the reference step) on the
    host: every usable core (OpenMP over tracers, at most 16 threads) on
    the whole step, then one core on the first eighth of the tracers
    (the tracer-independent part of a step is then a larger share, so that
    rate is a lower bound); about seconds / 2 each.  step(case, threads)
    runs one oracle step of one case."""
    model, ncpu, avail = cpu_info()
    ntr = max(c.ntr for c in cases)
    nth = max(1, min(16, avail))

    def sample(threads, m, budget):
        saved = [(c.wc.ntbdy, c.wc.ntdif_v) for c in cases]
        try:
            for c in cases:
                c.wc.ntbdy = min(c.wc.ntbdy, m)
                c.wc.ntdif_v = min(c.wc.ntdif_v, m)
            cu = float(sum(c.nwet * c.wc.ntbdy for c in cases))
            t = time.time()
            n = 0
            while True:
                for c in cases:
                    step(c, threads)
                n += 1
                if time.time() - t > budget or n >= 50:
                    break
            dt = time.time() - t
        finally:
            for c, (a, b) in zip(cases, saved):
                c.wc.ntbdy, c.wc.ntdif_v = a, b
        return cu * n / dt, n

    used = min(nth, ntr)
    v_all, n_all = sample(nth, ntr, seconds / 2)
    m1 = max(1, ntr // 8)
    v_one, n_one = sample(1, m1, seconds / 2)
    return {"value": v_all, "unit": "cell-updates/s", "cores": used, "kind": "port",
            "sample": "%d %s steps, all tracers, OpenMP over tracers on %d threads"
                      % (n_all, what, used),
            "single_core": {"value": v_one, "unit": "cell-updates/s", "cores": 1,
                            "sample": "%d %s steps of the first %d tracers, one thread"
                                      % (n_one, what, m1)},
            "nproc": ncpu, "usable_cpus": avail, "cpu_model": model}


def accuracy(got, ref, m_got, m_got0, m_ref, m_ref0, steps, floor=1e-3):
    """Error of a form against the FP64 plain form after `steps` steps.
    max_rel_err (the bar): max |got - ref| over the tracer's largest |ref|;
    max_pointwise_rel_err: max |got - ref| / |ref| over the cells holding at
    least floor times the tracer's largest |ref|; mass drift of the form
    against its own start and against the FP64 form's mass."""
    span = np.maximum(np.abs(ref).max(axis=1), 1e-30)
    err = np.abs(got - ref).max(axis=1) / span
    rms = np.sqrt(((got - ref) ** 2).mean(axis=1)) / span
    big = np.abs(ref) >= floor * span[:, None]
    pw = np.where(big, np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300), 0.0).max(axis=1)
    return {"steps": steps, "max_rel_err": float(err.max()),
            "max_pointwise_rel_err": float(pw.max()), "pointwise_floor": floor,
            "rms_rel_err": float(rms.max()),
            "mass_drift": float((np.abs(m_got - m_got0) / np.abs(m_got0)).max()),
            "mass_vs_fp64": float((np.abs(m_got - m_ref) / np.abs(m_ref)).max()),
            "mass_drift_fp64": float((np.abs(m_ref - m_ref0) / np.abs(m_ref0)).max()),
            "worst_tracer": int(np.argmax(err))}


def bench_ffsl(args):
    """The structured FFSL scheme (plain FP64 form) on the c3 basin."""
    from ems_amd import _lib, configs, structs as S
    from ems_amd import tracer as T
    L = _lib.lib()
    case = configs.c3(seed=args.seed, ntr=args.ntr)
    case.wc.trasc = S.FFSL
    case.wc.ultimate = 0
    case.wd.nstep = 1
    L.set_dz(case.gp, case.wp, case.pp)
    if _lib.last_error()[0]:
        raise RuntimeError("set_dz: %s" % (_lib.last_error(),))
    L.emstr_forget_window(case.gp)
    ev = HipEvents()
    K, W = min(args.steps, 20), min(args.warmup, 3)
    c = T.Context(case, T.PLAIN_F64)
    c.upload_fields()
    c.vel_w()
    c.upload()
    c.step(W)
    ms = timed_steps(c, ev, K) / K
    c.close()
    nwet, ntr = case.nwet, case.ntr
    cu = float(nwet) * ntr
    out = {"metric": "tracer cell-updates/s on 1 GPU (structured FFSL, FP64 plain)",
           "value": cu / (ms * 1e-3), "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded c3 generator, seed %d)" % args.seed,
           "config": {"workload": "c3 basin, FFSL: 400x600x40, %d tracers, dt=90 s" % ntr,
                      "n3_wet": nwet, "ntr": ntr},
           "step_hbm_fraction": (nwet * ntr * 16 + nwet * 56) / (ms * 1e-3) / 1e9 /
           HBM_PEAK_GBS}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_dz(case.gp, case.wp, case.pp)
        out["cpu_baseline"] = cpu_baselines([case], O.step, args.cpu_seconds, "c3 FFSL")
    print(json.dumps(out))


def bench_compas(args):
    """The COMPAS path on the C4 hexagon mesh: van Leer (FP64 plain form),
    or FFSL (streamlines, least squares, transverse terms) in the FP32 form
    beside the FP64 plain form, with the FP32 form's error and mass drift
    against the FP64 form after accuracy_steps (at most 500) steps."""
    from ems_amd import mesh as MS, structs_us as U
    from oracle import oracle_us as OU
    t0 = time.time()
    ffsl = args.workload == "c4-compas-ffsl"
    case = MS.c4(seed=args.seed, trasc=U.FFSL if ffsl else U.VANLEER)
    print("bench: c4 built in %.1f s: %d wet cells x %d tracers, %d wet edges"
          % (time.time() - t0, case.nwet, case.ntr, case.g.v3_e1), file=sys.stderr, flush=True)
    LU = MS.lib()
    LU.set_dz(case.gp, case.wp, case.pp)
    if MS.last_error()[0]:
        raise RuntimeError("set_dz: %s" % (MS.last_error(),))
    LU.emstr_us_forget_window(case.gp)
    case.w_from_continuity()
    ev = HipEvents()
    K, W = min(args.steps, 50), min(args.warmup, 5)
    nwet, ntr, ne = case.nwet, case.ntr, case.g.v3_e1
    cu = float(nwet) * ntr
    forms = [("plain", MS.PLAIN_F64)] + ([("fp32", MS.BATCH_F32)] if ffsl else [])
    ctxs, ms, m0 = {}, {}, {}
    flag0 = case.wc.flag
    for name, form in forms:
        # each context starts from the window's FFSL toggle state
        case.wc.flag = flag0
        c = MS.Context(case, form)
        c.upload()
        m0[name] = np.array([OU.tmass(case, n) for n in range(ntr)])
        c.step(W)
        c.sync()
        a, b = ev.event(), ev.event()
        ev.record(a, c.stream)
        c.step(K)
        ev.record(b, c.stream)
        ms[name] = ev.elapsed_ms(a, b) / K
        c.sync()
        ctxs[name] = (c, case.wc.flag)
        print("bench: c4 %s %.4f ms/step" % (name, ms[name]), file=sys.stderr, flush=True)
    head = "fp32" if ffsl else "plain"
    out = {"metric": "tracer cell-updates/s on 1 GPU (COMPAS, %s)"
                     % ("FFSL, FP32 form vs FP64 plain" if ffsl else "van Leer, FP64 plain"),
           "value": cu / (ms[head] * 1e-3), "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": ms[head], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32" if ffsl else "f64",
           "data": "synthetic (seeded C4 hexagon generator, seed %d)" % args.seed,
           "config": {"workload": "c4 mesh: 350x350 jittered hexagons, 1.5 km edges, 50 "
                                  "layers, 3-400 m, %d tracers, dt=180 s, %s"
                                  % (ntr, "FFSL (FP32 form: FP32 per-tracer FFSL work over "
                                     "FP64 maps, streamlines and state)"
                                     if ffsl else "van Leer"),
                      "n3_wet": nwet, "wet_edges": ne, "ntr": ntr},
           # per step: every tracer read and written once (FP64 state in both
           # forms), w, Kz, dz per cell and u1, u1flux3d per edge
           "step_hbm_fraction": (nwet * ntr * 16 + nwet * 24 + ne * 16) /
           (ms[head] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if ffsl:
        out["fp64_plain"] = {"value": cu / (ms["plain"] * 1e-3), "ms_per_step": ms["plain"]}
        out["ratio_fp32_over_fp64"] = ms["plain"] / ms["fp32"]
        # both carried on to the same step count from the same start
        S_ACC = max(W + K, min(args.accuracy_steps, 500))
        res = {}
        for name in ("plain", "fp32"):
            c, flag = ctxs[name]
            case.wc.flag = flag
            c.step(S_ACC - (W + K))
            c.sync()
            c.download()
            res[name] = (case.tracers(), np.array([OU.tmass(case, n) for n in range(ntr)]))
        out["accuracy"] = accuracy(res["fp32"][0], res["plain"][0], res["fp32"][1], m0["fp32"],
                                   res["plain"][1], m0["plain"], S_ACC)
    for c, _ in ctxs.values():
        c.close()
    LU.emstr_us_forget_window(case.gp)
    if not args.no_cpu:
        OU.set_dz(case.gp, case.wp, case.pp)
        out["cpu_baseline"] = cpu_baselines([case], OU.step, args.cpu_seconds, "c4 COMPAS")
    print(json.dumps(out))


def bench_c2(args):
    """The 200x300x30 curvilinear case (10 tracers, QUICKEST, dt = 120 s):
    the FP32 precision form against the FP64 plain form, rate of each and
    the precision form's error and mass drift after accuracy_steps (at
    most 500) steps."""
    from ems_amd import _lib, configs
    from ems_amd import tracer as T
    L = _lib.lib()
    case = configs.c2()
    L.set_dz(case.gp, case.wp, case.pp)
    if _lib.last_error()[0]:
        raise RuntimeError("set_dz: %s" % (_lib.last_error(),))
    L.emstr_forget_window(case.gp)
    ev = HipEvents()
    K, W = args.steps, args.warmup
    nwet, ntr = case.nwet, case.ntr
    cu = float(nwet) * ntr
    p = T.Context(case, T.PLAIN_F64)
    p.upload_fields()
    p.vel_w()
    p.upload()
    q = T.Context(case, T.PREC_F32)
    q.upload()
    mp0, mq0 = p.mass(), q.mass()
    res = {}
    for name, c in (("plain", p), ("precision", q)):
        c.step(W)
        res[name] = timed_steps(c, ev, K) / K
        print("bench: c2 %s %.4f ms/step" % (name, res[name]), file=sys.stderr, flush=True)
    S_ACC = max(W + K, min(args.accuracy_steps, 500))
    for c in (p, q):
        c.step(S_ACC - (W + K))
    mp, mq = p.mass(), q.mass()
    p.download()
    ref = case.tracers()
    q.download()
    acc = accuracy(case.tracers(), ref, mq, mq0, mp, mp0, S_ACC)
    p.close()
    q.close()
    ms = res["precision"]
    out = {"metric": "tracer cell-updates/s on 1 GPU (C2, FP32 precision form vs FP64 plain)",
           "value": cu / (ms * 1e-3), "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded c2 generator, seed 2)",
           "config": {"workload": "c2: 200x300x30 curvilinear basin, 30% land, 5-200 m, 10 "
                                  "tracers, QUICKEST, dt=120 s",
                      "n3_wet": nwet, "ntr": ntr, "form": "precision FP32 (FP64 budgets, "
                      "refinement pass, layer thicknesses)"},
           "fp64_plain": {"value": cu / (res["plain"] * 1e-3), "ms_per_step": res["plain"]},
           "ratio_fp32_over_fp64": res["plain"] / ms,
           "accuracy": acc}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_dz(case.gp, case.wp, case.pp)
        out["cpu_baseline"] = cpu_baselines([case], O.step, args.cpu_seconds, "c2")
    print(json.dumps(out))


def bench_ensemble(args):
    """The 64-window set c5 (van Leer, the batched form's scheme) in the
    batched FP32 form: one ensemble hipGraph with a branch per window,
    against the same windows stepped one context after another."""
    from ems_amd import _lib, configs, structs as S
    from ems_amd import tracer as T
    L = _lib.lib()
    t0 = time.time()
    doms = configs.c5(ndom=64, trasc=S.VANLEER)
    ctxs = []
    for c in doms:
        L.set_dz(c.gp, c.wp, c.pp)
        if _lib.last_error()[0]:
            raise RuntimeError("set_dz: %s" % (_lib.last_error(),))
        L.emstr_forget_window(c.gp)
        p = T.Context(c, T.PLAIN_F64)
        p.upload_fields()
        p.vel_w()
        p.close()
        b = T.Context(c, T.BATCH_F32)
        b.upload()
        ctxs.append(b)
    cu = float(sum(c.nwet * c.ntr for c in doms))
    print("bench: c5 built in %.1f s: %d windows, %.3g cell-tracers"
          % (time.time() - t0, len(doms), cu), file=sys.stderr, flush=True)
    K = max(2, min(args.steps, 100) // 2 * 2)
    # one window after another (each replays its own two-step graph)
    for b in ctxs:
        b.step(2)
        b.sync()
    t = time.time()
    for b in ctxs:
        b.step(K)
        b.sync()
    sep_ms = (time.time() - t) * 1e3 / K
    # per-window grids for the ensemble, by measurement
    tuned = None
    if not args.no_tune:
        from ems_amd import autotune
        bw, tms, trials = autotune.tune_ensemble(ctxs)
        tuned = {"block_width": bw, "ms_per_step": tms,
                 "trials": [[t[0], round(t[1], 4)] for t in trials]}
    # the ensemble graph
    ens = T.Ensemble(ctxs)
    ens.step(2)
    ens.sync()
    ev = HipEvents()
    a, e = ev.event(), ev.event()
    ev.record(a, ens.stream)
    ens.step(K)
    ev.record(e, ens.stream)
    ms = ev.elapsed_ms(a, e) / K
    ens.sync()
    ens.close()
    for b in ctxs:
        b.close()
    out = {"metric": "tracer cell-updates/s on 1 GPU (64-window ensemble, FP32 batched)",
           "value": cu / (ms * 1e-3), "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
           "warmup": 2, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded c5 generator, seed 5, van Leer)",
           "config": {"workload": "c5: 64 windows, nx, ny in [20, 150], 5-40 layers, 5-20 "
                                  "tracers, one ensemble hipGraph per two steps",
                      "windows": len(doms), "cell_tracers": cu},
           "one_after_another_ms_per_step": sep_ms,
           "ensemble_speedup": sep_ms / ms}
    if tuned:
        out["autotune"] = tuned
    # per-window accuracy: fresh FP64 plain contexts and a fresh ensemble of
    # batched contexts from the same initial state, S steps each
    S_ACC = max(2, min(args.accuracy_steps, 200) // 2 * 2)
    errs, pws, drifts, vs64 = [], [], [], []
    plains, fresh = [], []
    for c in doms:
        p = T.Context(c, T.PLAIN_F64)
        p.upload()
        plains.append(p)
        b = T.Context(c, T.BATCH_F32)
        b.upload()
        fresh.append(b)
    m0 = [b.mass() for b in fresh]
    ens = T.Ensemble(fresh)
    ens.step(S_ACC)
    ens.sync()
    ens.close()
    for c, p, b, mb0 in zip(doms, plains, fresh, m0):
        p.step(S_ACC)
        mp = p.mass()
        p.download()
        ref = c.tracers()
        mb = b.mass()
        b.download()
        a = accuracy(c.tracers(), ref, mb, mb0, mp, mp, S_ACC)
        errs.append(a["max_rel_err"])
        pws.append(a["max_pointwise_rel_err"])
        drifts.append(a["mass_drift"])
        vs64.append(a["mass_vs_fp64"])
        p.close()
        b.close()
    out["accuracy"] = {"steps": S_ACC, "max_rel_err": max(errs),
                       "max_pointwise_rel_err": max(pws), "pointwise_floor": 1e-3,
                       "mass_drift": max(drifts), "mass_vs_fp64": max(vs64),
                       "per_window": {"max_rel_err": [float("%.3g" % e) for e in errs],
                                      "mass_drift": [float("%.3g" % e) for e in drifts],
                                      "stiff": [bool(getattr(c, "stiff", False)) for c in doms]}}
    if not args.no_cpu:
        from oracle import oracle as O
        for c in doms:
            O.set_dz(c.gp, c.wp, c.pp)
        out["cpu_baseline"] = cpu_baselines(doms, O.step, args.cpu_seconds, "c5 (64 windows)")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
