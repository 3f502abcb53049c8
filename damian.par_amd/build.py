"""Build recipe for the step library and its companions.

    python damian.par_amd/build.py            # everything
Outputs (in-tree, git-ignored):
    damian.par_amd/lib/libdamian_hip.so       HIP kernels + C-ABI (gfx950)
    damian.par_amd/lib/damian_step.mod/.o     Fortran ISO_C_BINDING module (amdflang)
    damian.par_amd/lib/dmp_drive              Fortran driver linked to the library
    damian.par_amd/lib/dmp_dropin             drop-in init_explicit/run_explicit harness
    oracle/_build/liboracle.so                CPU oracle (test infrastructure)
    oracle/_ref/libexfem_io.so                reference io.cpp (only where the
                                              reference tree is present)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
CSRC = PKG / "csrc"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def run(cmd, **kw):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, **kw)


def build_hip():
    LIB.mkdir(exist_ok=True)
    out = LIB / "libdamian_hip.so"
    tmp = LIB / "libdamian_hip.so.tmp"
    run([HIPCC, *HIP_FLAGS, "-I", ROOT / "include", "-I", CSRC,
         CSRC / "dmp_kernels.hip", CSRC / "dmp_api.cpp", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_fortran():
    """ISO_C_BINDING module, the Fortran driver and the drop-in harness
    (INTEGRATION.md's init_explicit / run_explicit compiled against modules
    with the reference's declarations), linked against the library next to
    them (rpath $ORIGIN)."""
    F = PKG / "fortran"
    if not Path(FC).exists():
        print("fortran: skipped (amdflang not found)")
        return None
    LIB.mkdir(exist_ok=True)
    link = ["-L", LIB, "-ldamian_hip", "-Wl,-rpath,$ORIGIN"]
    run([FC, "-c", "-O2", "-J", LIB, F / "damian_step.f90", "-o", LIB / "damian_step.o"])
    run([FC, "-c", "-O2", "-J", LIB, F / "f1_site.f90", "-o", LIB / "f1_site.o"])
    run([FC, "-O2", "-I", LIB, "-J", LIB, F / "dmp_drive.f90", LIB / "damian_step.o",
         LIB / "f1_site.o", *link, "-o", LIB / "dmp_drive"])
    run([FC, "-O2", "-I", LIB, "-I", F, "-J", LIB, F / "dropin_harness.f90", LIB / "damian_step.o",
         LIB / "f1_site.o", *link, "-o", LIB / "dmp_dropin"])
    return LIB / "dmp_drive"


def build_oracle():
    run(["make", "-C", ROOT / "oracle", "-s"])


def build_all():
    build_oracle()
    build_hip()
    build_fortran()


if __name__ == "__main__":
    build_all()
