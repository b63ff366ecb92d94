"""Builds openradioss_amd/liborad.so (HIP, gfx950) in-tree with hipcc.

The FP64 element kernels and the nodal kernels are compiled without FMA
contraction so their arithmetic follows the reference's double-precision
build operation for operation; the mixed (FP32) element kernels keep
contraction.  Packed-FP32 instructions are disabled for every HIP file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liborad.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ORAD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-result", "-I", CSRC]
SOURCES = [
    ("orad_brick_fp64.hip", ["-ffp-contract=off"]),
    ("orad_brick_mixed.hip", ["-ffp-contract=fast"]),
    ("orad_shell_fp64.hip", ["-ffp-contract=off"]),
    ("orad_shell_mixed.hip", ["-ffp-contract=fast"]),
    ("orad_qeph_fp64.hip", ["-ffp-contract=off"]),
    ("orad_qeph_mixed.hip", ["-ffp-contract=fast"]),
    ("orad_qbat_fp64.hip", ["-ffp-contract=off"]),
    ("orad_qbat_mixed.hip", ["-ffp-contract=fast"]),
    ("orad_nodal.hip", ["-ffp-contract=off"]),
    ("orad_contact7.hip", ["-ffp-contract=off"]),
    ("orad_adjoint.hip", ["-ffp-contract=off"]),
    ("orad_brick_order.hip", []),
    ("orad_api.hip", []),
]


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [HIPCC, "-c"] + COMMON + extra + [os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs,
                   check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN_DIR, "resol_brick")
DRIVER_SHELL = os.path.join(FORTRAN_DIR, "resol_shell")


def build_fortran(verbose=False):
    """ISO_C_BINDING module + RESOL-style drivers (bricks, shells), linked against liborad.so."""
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    for drv in (DRIVER, DRIVER_SHELL):
        cmd = [fc, "-O2", "-ffp-contract=off", "-J", BUILD,
               os.path.join(FORTRAN_DIR, "orad_bind.F90"), drv + ".F90",
               "-L", HERE, "-lorad", "-Wl,-rpath,$ORIGIN/..", "-o", drv + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(drv + ".tmp", drv)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
