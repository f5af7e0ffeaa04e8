"""Build the native parts in-tree (gfx950 only).

    python -m qvort_amd.build          # libqvort_gpu.so + the Fortran driver

Outputs:
    qvort_amd/lib/libqvort_gpu.so      HIP kernels + C ABI (include/qvort_gpu.h)
    qvort_amd/lib/qv_run               Fortran driver (run.in + f(:) arrays)
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB = os.path.join(LIB_DIR, "libqvort_gpu.so")
DRIVER = os.path.join(LIB_DIR, "qv_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("QV_FC", "/opt/rocm/bin/amdflang")

# -ffp-contract=off: the FP64 path keeps the reference's operation order (no
# fused multiply-add), so its terms can match the reference bit for bit.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    # the ensemble kernel's 512-source tile shapes miss their 4-wave register
    # budget's occupancy on LDS, by design (qv_batch.hip)
    "-Wno-pass-failed",
]


def _run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lib(jobs: int = 8) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(LIB_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = []
    procs = []
    for src in srcs:
        obj = os.path.join(obj_dir, os.path.basename(src)[:-4] + ".o")
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str | None:
    srcs = [os.path.join(FSRC, f) for f in ("qv_gpu.f90", "qv_run.f90")]
    if not all(os.path.exists(s) for s in srcs) or not shutil.which(FC):
        return None
    mod_dir = os.path.join(LIB_DIR, "mod")
    os.makedirs(mod_dir, exist_ok=True)
    tmp = DRIVER + ".tmp"
    _run([FC, "-O2", "-module-dir", mod_dir, *srcs, "-o", tmp, "-L" + LIB_DIR, "-lqvort_gpu",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DRIVER)
    return DRIVER


def main() -> None:
    build_lib()
    build_driver()


if __name__ == "__main__":
    sys.exit(main())
