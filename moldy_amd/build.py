"""Builds the native parts in-tree.

* moldy_amd/lib/libmoldy_hip.so -- the HIP kernels + C ABI, for gfx950
* oracle/liboracle.so           -- the CPU oracle (test infrastructure)
* oracle/_ref/*                 -- the reference Fortran driver, only when the
                                   reference sources are present (build box)
* moldy_amd/lib/moldy_fortran   -- Fortran drop-in modules + driver (amdflang);
                                   lib/fec/ holds the -DIRON_CARBON build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MDK_OFFLOAD_ARCH", "gfx950")

# Packed FP32 VALU forms are disabled for all device code of this project.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wno-unused-value",
             "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_SOURCES = ["mdk_core.hip", "mdk_batch.hip"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{res.stdout[-4000:]}\n{res.stderr[-4000:]}")
    return res


def build_hip() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libmoldy_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_reference() -> bool:
    if not os.path.isdir("/root/reference/src"):
        return False
    _run([os.path.join(ROOT, "oracle", "build_ref.sh")])
    return True


def build_fortran() -> str | None:
    mk = os.path.join(HERE, "fortran", "Makefile")
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(mk) or not os.path.exists(fc):
        return None
    fdir = os.path.join(HERE, "fortran")
    _run(["make", "-C", fdir, "FC=" + fc])
    # the golden-vector driver, compiled unchanged against the drop-in modules
    _run(["make", "-C", fdir, "FC=" + fc, "dropin",
          "CALLER=" + os.path.join(ROOT, "oracle", "ref_driver.F90"),
          "OUT=" + os.path.join(LIBDIR, "ref_driver_dropin")])
    # the same with the iron_carbon material (-DIRON_CARBON, lib/fec/)
    _run(["make", "-C", fdir, "FC=" + fc, "fec"])
    _run(["make", "-C", fdir, "FC=" + fc, "dropin_fec",
          "CALLER=" + os.path.join(ROOT, "oracle", "ref_driver.F90"),
          "OUT=" + os.path.join(LIBDIR, "ref_driver_dropin_fec")])
    return os.path.join(LIBDIR, "libmoldy_fortran.a")


def build_all(verbose: bool = True) -> None:
    steps = [("hip", build_hip), ("oracle", build_oracle), ("reference", build_reference),
             ("fortran", build_fortran)]
    for name, fn in steps:
        r = fn()
        if verbose:
            print(f"[build] {name}: {r}", file=sys.stderr)


if __name__ == "__main__":
    build_all()
