"""The reference's own main program over the drop-in, built on the CPU
(moldy_amd/fortran/build_refmain.sh): src/moldin.F90 with its io_m,
analysis_m and metropolis_m, and the reference's state modules (params_m,
particles_m, system_m, ...), compiled unchanged around the drop-in's hot-path
modules (potential_m, linkcell_m, neighbourlist_m, thermostat_m, dynamics_m,
quench_m) and linked against libmoldy_hip.so.  Every symbol the program needs
resolves except rand_: metropolis_m calls RAND, an intrinsic of the GNU and
PGI Fortran runtimes that the reference's Makefile targets and amdflang's
runtime lacks, so the program links with those compilers only.

Needs the reference tree (the test reads its sources where they lie) and
amdflang; skipped where either is absent."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

REF_SRC = "/root/reference/src"
SCRIPT = os.path.join(ROOT, "moldy_amd", "fortran", "build_refmain.sh")
LIB = os.path.join(ROOT, "moldy_amd", "lib", "libmoldy_hip.so")


def test_reference_main_program_links_against_the_dropin(tmp_path):
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    if not os.path.isdir(REF_SRC) or not (os.path.exists(fc) or shutil.which(fc)):
        pytest.skip("reference sources or amdflang not available")
    if not os.path.exists(LIB):
        pytest.skip("libmoldy_hip.so not built")
    res = subprocess.run(["bash", SCRIPT, REF_SRC, str(tmp_path)], capture_output=True, text=True,
                         timeout=600)
    undefined = sorted(ln.split(":", 1)[1].strip() for ln in res.stdout.splitlines()
                       if ln.startswith("undefined symbol:"))
    assert undefined == ["rand_"], (undefined, res.stderr[-2000:])
    for obj in ("moldin.o", "io_m.o", "analysis_m.o", "metropolis_m.o", "hotpath_m.o"):
        assert (tmp_path / obj).exists(), obj
