"""Build the native libraries in-tree.

  fooskagpu_amd/lib/libfooska_mesh.so  host mesh/case library (g++)
  fooskagpu_amd/lib/libfooska_hip.so   gfx950 kernels + C ABI (hipcc)
  oracle/liboracle{,_omp}.so           CPU FP64 oracle (make -C oracle)

Usage: python -m fooskagpu_amd.build [mesh|hip|oracle ...]
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FK_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no implicit FMA, so the FP64 path reproduces the serial
# FP64 evaluation of the same expressions bit for bit.
# -packed-fp32-ops: keep the compiler from emitting packed FP32 arithmetic.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 divide and sqrt as reciprocal / hardware sqrt (about 1 ulp) instead
    # of the correctly rounded expansions: only the mixed path and the MLP
    # compute in FP32; the FP64 path is unaffected
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_mesh():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfooska_mesh.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          os.path.join(CSRC, "mesh.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfooska_hip.so")
    srcs = [os.path.join(CSRC, "fk_kernels.hip"), os.path.join(CSRC, "fk_mlp.hip")]
    _run([HIPCC] + HIP_FLAGS + srcs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "all"])


def build_all():
    build_mesh()
    build_oracle()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["mesh", "oracle", "hip"]
    for w in what:
        {"mesh": build_mesh, "hip": build_hip, "oracle": build_oracle}[w]()
