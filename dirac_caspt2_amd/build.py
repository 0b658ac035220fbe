"""Build libdcaspt2_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m dirac_caspt2_amd.build

Every HIP source is compiled for gfx950 only, with packed-FP32 instructions
disabled, and linked against rocSOLVER/rocBLAS (eigensolvers and LU).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdcaspt2_amd.so")
SOURCES = ["dc2_gemm.hip", "dc2_ozaki.hip", "dc2_ci.hip", "dc2_pt2.hip", "dc2_ingest.hip", "dc2_heev.hip", "dc2_tridc.hip",
           "dc2_api.hip",
           "dc2_serve.hip", "dc2_noalloc.hip"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
FORTRAN_SRC = os.path.join(HERE, "fortran", "dcaspt2_amd_interface.F90")


def _run(cmd, quiet=True):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(ln for ln in res.stdout.splitlines() if "packed-fp32-ops" not in ln)
    if res.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    if out.strip() and not quiet:
        print(out)
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c", os.path.join(CSRC, src), "-o", obj],
             quiet=not verbose)
        return obj

    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L", os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas", "-Wl,-rpath," + os.path.join(ROCM, "lib")],
         quiet=not verbose)
    os.replace(tmp, LIB)
    build_fortran_interface()
    return LIB


FORTRAN_DRIVER_SRC = os.path.join(HERE, "..", "tests", "fortran", "drive_caspt2.F90")
FORTRAN_DRIVER = os.path.join(BUILD, "drive_caspt2")


def build_fortran_interface() -> None:
    """Compile the iso_c_binding interface module with amdflang and link the Fortran
    test driver against libdcaspt2_amd.so (a Fortran host calling the C ABI)."""
    flang = shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang")
    if not os.path.exists(FORTRAN_SRC) or not os.path.exists(flang):
        return
    obj = os.path.join(BUILD, "dcaspt2_amd_interface.o")
    _run([flang, "-O2", "-c", FORTRAN_SRC, "-module-dir", BUILD, "-o", obj])
    if os.path.exists(FORTRAN_DRIVER_SRC):
        llvm_lib = os.path.join(ROCM, "llvm", "lib")
        _run([flang, "-O2", "-module-dir", BUILD, "-I", BUILD, FORTRAN_DRIVER_SRC, obj, "-o", FORTRAN_DRIVER + ".tmp",
              "-L", HERE, "-ldcaspt2_amd", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + llvm_lib,
              "-Wl,-rpath," + os.path.join(ROCM, "lib")])
        os.replace(FORTRAN_DRIVER + ".tmp", FORTRAN_DRIVER)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
