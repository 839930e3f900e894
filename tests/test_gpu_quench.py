"""The conjugate-gradient quench on the device force (moldy_amd.quench over
the C ABI): against the reference's own quench (VA14CD) on a 249-atom vacancy
cell -- energy within the reference's stopping tolerance 5e-7 n eV, positions
2e-4 fractional, box 1e-4 -- and the perfect lattice's closed-form
equilibrium lattice parameter at 0 and 5 GPa (2e-6 relative)."""
import pytest

from conftest import golden_system, load_golden, run_kwargs
from moldy_amd import engine
import quench_case as qc

pytestmark = pytest.mark.gpu


def test_gpu_quench_matches_reference_vacancy():
    F = load_golden("quench_vac_249")
    print("[gpu-test] quench, vacancy cell", flush=True)
    eng = engine.Engine(golden_system(F), **run_kwargs(F))
    eng.linkup()
    eng.md_init()
    qc.check_vacancy(eng, F)
    eng.close()


def test_gpu_quench_perfect_lattice_closed_form(oracle):
    print("[gpu-test] quench, perfect lattice", flush=True)
    engines = []

    def make(sd, press):
        e = engine.Engine(sd, ivol=0, press_gpa=press)
        e.linkup()
        engines.append(e)
        return e

    for press in (0.0, 5.0):
        qc.check_perfect_lattice(make, oracle.Potential(), press)
    for e in engines:
        e.close()
