"""Compile the HIP sources in pyclaw_amd/csrc into pyclaw_amd/libclawhip.so (gfx950).

Flags:
  -ffp-contract=off   evaluation order and rounding of every FP64 expression
                      match the reference's Fortran (no FMA contraction)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from emitting packed FP32 arithmetic
  -mllvm -amdgpu-mfma-vgpr-form
                      MFMA results in VGPRs (no accumulator-register copies)
  -fno-hip-fp32-correctly-rounded-divide-sqrt
                      FP32 division and square root by the hardware reciprocal
                      and square-root instructions (the FP32 paths are checked
                      against tolerances; FP64 division and sqrt stay IEEE)
  -fgpu-flush-denormals-to-zero
                      FP32 denormals flushed (FP64 denormals are kept): the FP32
                      reciprocal and square root then need no range scaling
                      (about 8 % fewer VALU instructions in the FP32 stage march);
                      the FP32 states are O(1), denormals appear only in exact
                      cancellation's neighbourhood

The kernel sources in VARIANT_SRCS are compiled twice: as they are (namespace
claw::exact, no contraction: CLAWHIP_F64) and with -DCLAW_FAST=1 (claw::fast,
a*b + c in one expression contracted to an FMA: CLAWHIP_F32K and
CLAWHIP_F64_FAST); csrc/variant.hpp describes the two.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libclawhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CLAWHIP_ARCH", "gfx950")

FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
         "-fgpu-flush-denormals-to-zero",
         f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function"]


VARIANT_SRCS = ("classic2.hip", "classic_split.hip", "ensemble.hip", "sharp2.hip")


def build(verbose=False, jobs=8):
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    units = [(s, os.path.basename(s) + ".o", []) for s in srcs]
    units += [(s, os.path.basename(s) + ".fast.o", ["-DCLAW_FAST=1"]) for s in srcs
              if os.path.basename(s) in VARIANT_SRCS]
    # the largest sources first, so the parallel compile ends together
    units.sort(key=lambda u: -os.path.getsize(u[0]))
    procs = []
    objs = []
    for s, name, extra in units:
        o = os.path.join(objdir, name)
        objs.append(o)
        cmd = [HIPCC] + FLAGS + extra + ["-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    while procs:
        _wait(procs.pop(0))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if out.returncode != 0:
        raise RuntimeError("link failed:\n" + out.stdout.decode())
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out.decode()}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
