"""Build libphsph.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m phantom_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libphsph.so")
SOURCES = ["api.hip", "tree.hip", "density.hip", "force.hip", "gravity.hip", "step.hip", "substep.hip", "vjp.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PHSPH_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


FORTRAN = os.path.join(HERE, "fortran")
FC = os.environ.get("FC_PHSPH", "/opt/rocm/bin/amdflang")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose=False):
    """The iso_c_binding module and the Fortran drivers that call the
    library the way Phantom's deriv.f90 and step do (skipped without
    amdflang)."""
    if not os.path.exists(FC):
        if verbose:
            print("amdflang not found: Fortran interface not built", flush=True)
        return None
    mod = os.path.join(FORTRAN, "phsph.F90")
    obj = os.path.join(FORTRAN, "phsph.o")
    cmds = [[FC, "-O2", "-c", mod, "-J", FORTRAN, "-o", obj]]
    for name in ("shim_driver", "step_driver"):
        drv = os.path.join(FORTRAN, name + ".f90")
        exe = os.path.join(FORTRAN, name)
        cmds.append([FC, "-O2", "-I", FORTRAN, "-J", FORTRAN, drv, obj, "-L", HERE, "-lphsph",
                     "-Wl,-rpath,$ORIGIN/..", "-o", exe])
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        subprocess.run(c, check=True)
    return exe


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
