"""In-tree build of the native libraries (no JIT cache, nothing installed).

    python -m tmcore_amd.build            # libtmhost.so + libtmgpu.so (+ Fortran driver)

libtmhost.so  g++     mesh generator, poly_fit weights, initial states
libtmgpu.so   hipcc   gfx950 kernels + C-ABI (include/tmcore_gpu.h)
test_tmcore_sw amdflang  Fortran namelist driver over the C-ABI (if amdflang exists)
libtmcpu.so   amdflang  Fortran FP64 CPU path (OpenMP), the CPU baseline of bench.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for every device compile.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-shared", "-std=c++17",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INC, "-I" + os.path.join(CSRC, "gpu")]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_host():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, "host", f) for f in ("mesh_icos.cpp", "mpas_grid.cpp", "poly_fit.cpp", "reconstruct.cpp", "history.cpp", "cases.cpp")]
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I" + INC, *srcs,
          "-o", os.path.join(LIB, "libtmhost.so")])


def build_gpu():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, "gpu", f) for f in ("tmgpu.hip", "ens.hip", "mlp.hip")]
    _run([HIPCC, *HIP_FLAGS, *srcs, "-o", os.path.join(LIB, "libtmgpu.so")])


def build_fortran():
    """Fortran host driver test_tmcore_sw (iso_c_binding over the C-ABI, or
    the host CPU path with use_gpu = .false.), the operator-wrapper test
    program, and libtmcpu.so (the Fortran FP64 CPU path alone)."""
    if not os.path.exists(FLANG):
        print("amdflang not found; skipping the Fortran driver", flush=True)
        return
    fdir = os.path.join(PKG, "fortran")
    srcs = ["tm_desc_mod.F90", "tm_c_binding_mod.F90", "tm_params_mod.F90", "tm_swm_cpu_mod.F90", "tm_swm_gpu_mod.F90",
            "test_tmcore_sw.F90"]
    if not all(os.path.exists(os.path.join(fdir, s)) for s in srcs):
        return
    bdir = os.path.join(fdir, "build")
    os.makedirs(bdir, exist_ok=True)
    link = ["-L" + LIB, "-ltmgpu", "-ltmhost", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    # -fopenmp and no FMA contraction for the CPU path the host can run (use_gpu = .false.)
    fflags = ["-O3", "-fopenmp", "-ffp-contract=off"]
    _run([FLANG, *fflags, "-J", bdir, *[os.path.join(fdir, s) for s in srcs], "-o",
          os.path.join(LIB, "test_tmcore_sw"), *link], cwd=bdir)
    # the operator wrappers under their reference names, for the parity test
    _run([FLANG, *fflags, "-J", bdir, *[os.path.join(fdir, s) for s in srcs[:-1]],
          os.path.join(fdir, "test_tmcore_ops.F90"), "-o", os.path.join(LIB, "test_tmcore_ops"), *link], cwd=bdir)
    # the FP64 CPU path: -O3 with OpenMP, no FMA contraction (the reference's
    # association order, as the oracle keeps it)
    cdir = os.path.join(bdir, "cpu")
    os.makedirs(cdir, exist_ok=True)
    _run([FLANG, "-O3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-J", cdir,
          os.path.join(fdir, "tm_desc_mod.F90"), os.path.join(fdir, "tm_swm_cpu_mod.F90"),
          "-o", os.path.join(LIB, "libtmcpu.so")], cwd=cdir)


def build_all():
    build_host()
    build_gpu()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "host": build_host, "gpu": build_gpu, "fortran": build_fortran}[w]()
