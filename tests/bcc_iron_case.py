"""The reference's own shipped test case tests/reference/bcc_iron (helper, not
collected): 31,250 Fe atoms (25^3 bcc cells, a = 2.870 A), iron_carbon
potential through 5000-point lookup tables, Parrinello-Rahman NPT at
0.1 GPa, 600 K velocities from setvel, velocity Verlet, 10 steps.

The fixture tests/golden/bcc_iron_golden.npz holds what the reference printed
(output66.txt, system.out) with half a unit of each last printed digit.  The
initial state is rebuilt as moldin makes it (RANMAR, read_system's
replication, setvel) by the oracle's input generator.
"""
from __future__ import annotations

import numpy as np

# eV/A^3 -> GPa factor printed by runavs (src/analysis_m.F90:301)
EV_A3_TO_GPA = 160.2176487


def load():
    from conftest import load_golden
    return load_golden("bcc_iron_golden")


def initial_state(F):
    """lattice.System of the shipped test's first step (positions and setvel velocities)."""
    import pyoracle
    from moldy_amd import lattice
    b0, s, x1, z, mass = pyoracle.moldin_new_run(F["basis"], F["basis_z"], F["basis_mass"], F["nt"],
                                                 F["cell"], float(F["dsp"]), float(F["temprq"]),
                                                 float(F["deltat_fs"]))
    return lattice.System(b0=b0, s=s, x1=x1, mass=mass, z=z, deltat_fs=float(F["deltat_fs"]))


def summarize(steps, tp, b0, press_model, nm, bk):
    """runavs quantities (src/analysis_m.F90:237-354 and moldin.F90:752-791).

    steps: per step (pe, ke, vol, box kinetic energy, b0 (3x3)) after each MD
    step; tp, b0: tensors after the last step.  The shipped output66.txt comes
    from an older runavs whose STRESS block printed -B0 tp / vol (the virial
    alone), which is what is computed here."""
    pe = np.array([r[0] for r in steps])
    ke = np.array([r[1] for r in steps])
    vol = np.array([r[2] for r in steps])
    bke = np.array([r[3] for r in steps])
    te = ke + pe
    h = te + press_model * vol
    th = h + bke
    out = {}

    def put(key, x):
        # update_thermodynamic_sums / runavs (src/analysis_m.F90:142-157, 248-263):
        # sequential sums; the spreads are differences of nearly equal numbers,
        # so the reference's own arithmetic is kept
        s = ssq = 0.0
        for v in x:
            s = s + float(v)
            ssq = ssq + float(v) * float(v)
        rt = 1.0 / len(x)
        a = rt * s
        out[key] = x[-1]
        out["a" + key] = a
        out["d" + key] = np.sqrt(abs(rt * ssq - a * a))

    for key, x in (("pe", pe), ("ke", ke), ("te", te), ("h", h), ("th", th), ("vol", vol)):
        put(key, x)
    out["t"] = out["ke"] / (1.5 * bk * nm)       # set_temp(ke/(1.5 bk nm)), moldin.F90:763
    out["at"] = out["ake"] / (1.5 * nm * bk)     # runavs :259-260
    out["dt"] = out["dke"] / (1.5 * nm * bk)
    out["stress"] = -(np.asarray(b0) @ np.asarray(tp)) / vol[-1] * EV_A3_TO_GPA
    sb0 = np.zeros((3, 3))
    for r in steps:
        sb0 = sb0 + np.asarray(r[4])
    out["b0_mean"] = sb0 * (1.0 / len(steps))
    out["b0"] = np.asarray(b0)
    return out


def check(F, got, offdiag_rtol=None):
    """Compares runavs quantities with the printed ones: each within half a
    unit of its last printed digit (plus 1e-12 relative).  offdiag_rtol, when
    given, replaces that bound for the box's and the stress's off-diagonal
    elements, which are tiny differences of large cancelling sums."""
    bad = []
    for key in ("pe", "ape", "dpe", "ke", "ake", "dke", "te", "ate", "dte", "h", "ah", "dh", "th", "ath",
                "dth", "vol", "avol", "dvol", "t", "at", "dt"):
        want, tol = float(F["o66_" + key]), float(F["o66_" + key + "_tol"])
        if not abs(got[key] - want) <= tol * 1.0001 + 1e-12 * abs(want):
            bad.append(f"{key}: {got[key]!r} vs printed {want!r} (+/- {tol})")
    off = ~np.eye(3, dtype=bool)
    for key in ("stress", "b0_mean", "b0"):
        want, tol = np.asarray(F["o66_" + key]), np.asarray(F["o66_" + key + "_tol"])
        g = np.asarray(got[key])
        ok = np.abs(g - want) <= tol * 1.0001 + 1e-12 * np.abs(want)
        if offdiag_rtol is not None:
            ok[off] = np.abs(g - want)[off] <= offdiag_rtol * np.abs(want)[off] + tol[off]
        if not ok.all():
            bad.append(f"{key}: {g.ravel()} vs printed {want.ravel()}")
    return bad


def check_system_out(F, b0, x0, y0, z0, en, box_rtol=0.0, offdiag_rtol=None):
    """system.out: full-precision box, and the sampled atoms' fractional
    positions (6 significant digits) and energies (4 decimals)."""
    bad = []
    want = np.asarray(F["sys_b0"])
    diff = np.abs(np.asarray(b0) - want)
    lim = box_rtol * np.abs(want) + 4e-15 * np.abs(want).max()
    if offdiag_rtol is not None:
        off = ~np.eye(3, dtype=bool)
        lim[off] = offdiag_rtol * np.abs(want)[off]
    if not (diff <= lim).all():
        bad.append(f"box: {np.asarray(b0).ravel()} vs {want.ravel()}")
    idx = np.asarray(F["sys_index"])
    for nm, arr in (("x0", x0), ("y0", y0), ("z0", z0), ("en", en)):
        w, t = np.asarray(F["sys_" + nm]), np.asarray(F["sys_" + nm + "_tol"])
        g = np.asarray(arr)[idx]
        # a printed 6-digit value is the true value rounded: allow the half unit
        # and a hair for values that sit on a rounding boundary
        k = np.abs(g - w) > t * 1.0001 + 1e-12
        if k.any():
            i = int(np.argmax(np.abs(g - w) - t))
            bad.append(f"{nm}: {int(k.sum())} sampled atoms off, e.g. atom {int(idx[i])}: {g[i]!r} vs {w[i]!r}")
    return bad
