"""Build libpam.so (HIP, gfx950) in-tree with hipcc.

``python -m pseudoacousticmodeling_amd.build`` compiles the kernels and the C
ABI into ``pseudoacousticmodeling_amd/libpam.so`` and the C++ drivers
``run_modeling`` / ``run_migration`` next to it.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpam.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PAM_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code: the compiler may
# form v_pk_* from ordinary float code, and some of their forms have been
# unreliable on the target machines.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                # the z-march keeps its row queues in registers only when the
                # row loop is fully unrolled; the default threshold gives up on
                # the 4-cells-per-lane variants and spills the queues to scratch
                "-mllvm", "-pragma-unroll-threshold=500000",
                # products fuse into additions only within one expression, as
                # written: under the default (fast) contraction a product fuses
                # across statements depending on the basic-block layout of each
                # instance, so two kernels instantiating the same step body
                # (a grid alone, a group of grids; the forward-only and adjoint
                # instances) could round one cell differently
                "-ffp-contract=on"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", f"-I{os.path.join(ROOT, 'include')}"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(verbose: bool = True) -> str:
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    srcs = [(f, True) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")] + [("pam_api.cpp", False)]
    procs = []
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    for name, dev in srcs:
        while len([p for p in procs if p.poll() is None]) >= jobs:
            time.sleep(0.2)
        obj = os.path.join(HERE, "_obj", name + ".o")
        cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, "-c"]
        if dev:
            cmd += ["-x", "hip"]
        cmd += [os.path.join(CSRC, name), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_cli(verbose: bool = True) -> list:
    """The C++ drivers run_modeling / run_migration (csrc/cli), linked against
    libpam.so next to them (rpath $ORIGIN/..)."""
    out = []
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    for name in ("run_modeling", "run_migration"):
        exe = os.path.join(bindir, name)
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
               os.path.join(CSRC, "cli", name + ".cpp"), "-o", exe + ".tmp", f"-L{HERE}", "-lpam",
               "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def main(argv=None) -> int:
    build_lib()
    build_cli()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
