"""Builds the HIP library librmpcd_amd.so in-tree for gfx950 with hipcc.

    python -m rmpcdmd_amd.build            # or rmpcdmd_amd.build.build()
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librmpcd_amd.so")
SOURCES = ["rmpcd_runtime.hip", "rmpcd_sort.hip", "rmpcd_mpcd.hip", "rmpcd_md.hip",
           "rmpcd_nlist.hip", "rmpcd_sim.hip", "rmpcd_fused.hip", "rmpcd_programs.hip",
           "rmpcd_bodies.hip", "rmpcd_colloids.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(REPO, "include"), "-I", CSRC, "-Wall", "-Wno-unused-function"]


def _compile(src):
    out = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return out


FORTRAN = os.path.join(HERE, "fortran")
FBIN = os.path.join(FORTRAN, "bin")
FLIB = os.path.join(FORTRAN, "librmpcd_fortran.so")
AMDFLANG = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")
FPROGRAMS = ["fortran_cells", "fortran_fluid", "fortran_dimer", "fortran_walls", "fortran_bodies",
             "fortran_params", "single_dimer_pbc", "fortran_shims"]


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s%s" % (what, r.stdout, r.stderr))


def build_fortran(verbose=False):
    """the Fortran module API (amdflang) over librmpcd_amd.so, and its driver programs"""
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    os.makedirs(FBIN, exist_ok=True)
    objs = []
    for src in ("rmpcd_capi.f90", "shims/hdf5.f90", "shims/h5md_module.f90", "shims/tester.f90", "rmpcdmd.f90",
                "particle_system_io.f90", "rmpcdmd_module.f90", "parsetext.f90"):
        out = os.path.join(OBJ, os.path.basename(src).replace(".f90", ".o"))
        _run([AMDFLANG, "-c", "-O2", "-fPIC", "-J", mod, os.path.join(FORTRAN, src), "-o", out], src)
        objs.append(out)
    _run([AMDFLANG, "-shared", "-fPIC", "-o", FLIB, *objs, "-L" + HERE, "-lrmpcd_amd",
          "-Wl,-rpath,$ORIGIN/.."], "librmpcd_fortran.so")
    for prog in FPROGRAMS:
        _run([AMDFLANG, "-O2", "-J", mod, "-I", mod, os.path.join(FORTRAN, "programs", prog + ".f90"),
              "-o", os.path.join(FBIN, prog), "-L" + FORTRAN, "-lrmpcd_fortran", "-L" + HERE, "-lrmpcd_amd",
              "-Wl,-rpath,$ORIGIN/..:$ORIGIN/../.."], prog)
    if verbose:
        print("built", FLIB, "and", ", ".join(FPROGRAMS))


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
