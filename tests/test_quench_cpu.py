"""The conjugate-gradient quench (moldy_amd.quench) driven by the oracle's
force: against the reference's own quench (VA14CD) on a vacancy cell, and the
perfect lattice's closed-form equilibrium lattice parameter."""
import numpy as np

from conftest import golden_system, load_golden, run_kwargs
import quench_case as qc


class OracleEngine:
    """The engine interface quench() uses, over an oracle system (the list is
    rebuilt at every force evaluation)."""

    def __init__(self, oracle, S):
        self.oracle, self.S, self.n = oracle, S, S.n

    def upload(self, nm, v):
        self.S.set_array(nm, v)

    def set_tensor(self, nm, t):
        self.S.set_tensor(nm, t)

    def force(self):
        self.oracle.lib().orc_set_ntc(self.S.h, 0)
        assert self.S.call("force") == 0

    def scalars(self):
        return {"pe": self.S.scalar("pe"), "vol": self.S.scalar("vol")}

    def array(self, nm):
        return self.S.array(nm)

    def tensor(self, nm):
        return self.S.tensor(nm)


def test_quench_matches_reference_vacancy(oracle):
    F = load_golden("quench_vac_249")
    S = oracle.OracleSystem(golden_system(F), oracle.Potential(), **run_kwargs(F))
    S.call("linkup")
    S.call("md_init")
    qc.check_vacancy(OracleEngine(oracle, S), F)


def test_quench_perfect_lattice_closed_form(oracle):
    pot = oracle.Potential()

    def make(sd, press):
        S = oracle.OracleSystem(sd, pot, ivol=0, press_gpa=press)
        S.call("linkup")
        return OracleEngine(oracle, S)

    for press in (0.0, 5.0):
        qc.check_perfect_lattice(make, pot, press, nc=4)
