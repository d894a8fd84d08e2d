"""Builds libemstr.so (HIP kernels for gfx950 + the C host layer) in-tree.

    python -m ems_amd.build

HIP sources are compiled with hipcc for gfx950 only, without FMA
contraction (the plain form reproduces the reference's FP64 operation
order; the batched FP32 form alone may contract) and without packed-FP32
instructions; the C sources with gcc.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SRC = ["emstr_plain.hip", "emstr_prec.hip", "emstr_batch.hip", "emstr_scratch.hip"]
C_SRC = ["emstr_host.c", "emstr_grid.c", "emstr_err.c"]

HIPFLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-Wno-unused-result"]
# the batched FP32 form is not bound to the reference's operation order:
# let the compiler fuse multiply-adds there
HIP_EXTRA = {"emstr_batch.hip": ["-ffp-contract=fast"]}
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
          "-I" + os.path.join(ROCM, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % cmd[-1])
    return r


# libemstr_us.so: the COMPAS path (unstructured meshes)
US_HIP_SRC = ["emstr_us.hip", "emstr_us_ffsl.hip"]
US_C_SRC = ["emstr_mesh.c", "emstr_lsq.c", "emstr_us_host.c", "emstr_err.c"]
# the column kernels of the plain form (set_dz, vertical diffusion, ...)
US_SHARED_HIP = ["emstr_plain.hip"]


def build_us(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    for s in US_HIP_SRC:
        o = os.path.join(BUILD, "us_" + s.replace(".hip", ".o"))
        _run([hipcc] + HIPFLAGS + HIP_EXTRA.get(s, []) + INC +
             ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in US_SHARED_HIP:
        # compiled again here (never an object left by build()), so that
        # the COMPAS library always links the current source
        o = os.path.join(BUILD, "us_" + s.replace(".hip", ".o"))
        _run([hipcc] + HIPFLAGS + HIP_EXTRA.get(s, []) + INC +
             ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in US_C_SRC:
        o = os.path.join(BUILD, "us_" + s.replace(".c", ".o"))
        _run(["gcc"] + CFLAGS + INC + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    out = os.path.join(HERE, "libemstr_us.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread", "-lm"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    for s in HIP_SRC:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        _run([hipcc] + HIPFLAGS + HIP_EXTRA.get(s, []) + INC +
             ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in C_SRC:
        o = os.path.join(BUILD, s.replace(".c", ".o"))
        _run(["gcc"] + CFLAGS + INC + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    out = os.path.join(HERE, "libemstr.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread", "-lm"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_us(verbose=True)
