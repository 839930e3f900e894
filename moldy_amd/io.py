"""Text outputs of a MOLDY run: system.out and the running-averages block.

write_system_out follows write_system_out_file (src/io_m.F90:363-379): the
atom count, "1 1 1", the three cell vectors B0(:,i), then one line per atom
in FORMAT(3g14.6,1X,I3,g12.4,f13.9) -- fractional x, y, z, atomic number,
mass, per-atom energy (energy_calc).

ThermoSums / runavs_text follow update_thermodynamic_sums and runavs
(src/analysis_m.F90:142-157, 237-354): sums of pe, ke, te, h, th, vol, their
squares and of B0 over the sampled steps; averages, spreads
sqrt(|<x^2> - <x>^2|), temperatures, the TP+TK stress in GPa and the mean and
instantaneous box, in runavs's FORMAT statements 1, 2, 3, 99, 6 and 8.
Fortran edit descriptors are rendered by fortran_e / fortran_f / fortran_g
(rounding half away from zero on the decimal digits as Python formats them).
"""
from __future__ import annotations

import math

import numpy as np

EV_A3_TO_GPA = 160.2176487   # src/analysis_m.F90:301


def fortran_e(x: float, w: int, d: int, e: int = 2) -> str:
    """Ew.d (Ee): 0.ddddE+xx, right-justified in w columns."""
    if x == 0.0 or not math.isfinite(x):
        body = f"0.{'0' * d}E+{'0' * e}" if x == 0.0 else str(x)
        return body.rjust(w)
    m = f"{abs(x):.{d - 1}e}"          # d.ddd e+xx: one more leading digit than Fortran
    mant, ex = m.split("e")
    digits = mant.replace(".", "")
    exp = int(ex) + 1
    s = f"{'-' if x < 0 else ''}0.{digits}E{'+' if exp >= 0 else '-'}{abs(exp):0{e}d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_f(x: float, w: int, d: int) -> str:
    s = f"{x:.{d}f}" + ("." if d == 0 else "")
    if s.startswith("0.") and len(s) > w:
        s = s[1:]
    elif s.startswith("-0.") and len(s) > w:
        s = "-" + s[2:]
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_g(x: float, w: int, d: int) -> str:
    """Gw.d for reals: F editing with d significant digits and 4 trailing
    blanks when 0.1 <= |x| < 10**d (after rounding), E editing otherwise."""
    a = abs(x)
    if a == 0.0:
        return fortran_f(0.0, w - 4, d - 1) + "    "
    if 0.1 - 0.5 * 10.0 ** (-d - 1) <= a < 10.0 ** d - 0.5:
        k = int(math.floor(math.log10(float(f"{a:.{d - 1}e}")))) + 1   # digits before the point
        return fortran_f(x, w - 4, d - k) + "    "
    return fortran_e(x, w, d)


def write_system_out(path: str, b0, x0, y0, z0, atomic_number, mass, en_atom) -> None:
    """system.out as write_system_out_file writes it (src/io_m.F90:363-379)."""
    b0 = np.asarray(b0, dtype=np.float64)
    n = len(x0)
    with open(path, "w") as fh:
        fh.write(f"{n:12d}\n")
        fh.write(" 1 1 1\n")
        for i in range(3):
            fh.write("".join(f"{b0[j, i]:25.17E}" for j in range(3)) + "\n")
        for i in range(n):
            fh.write(fortran_g(x0[i], 14, 6) + fortran_g(y0[i], 14, 6) + fortran_g(z0[i], 14, 6) + " " +
                     f"{int(atomic_number[i]):3d}" + fortran_g(mass[i], 12, 4) + fortran_f(en_atom[i], 13, 9) +
                     "\n")


class ThermoSums:
    """update_thermodynamic_sums (src/analysis_m.F90:142-157)."""

    KEYS = ("pe", "te", "ke", "h", "th", "vol")

    def __init__(self):
        self.s = {k: 0.0 for k in self.KEYS}
        self.sq = {k: 0.0 for k in self.KEYS}
        self.sb0 = np.zeros((3, 3))
        self.numsum = 0
        self.last = {}

    def update(self, pe: float, ke: float, te: float, h: float, th: float, vol: float, b0) -> None:
        v = dict(pe=pe, ke=ke, te=te, h=h, th=th, vol=vol)
        for k in self.KEYS:
            self.s[k] = self.s[k] + v[k]
            self.sq[k] = self.sq[k] + v[k] * v[k]
        self.sb0 = self.sb0 + np.asarray(b0, dtype=np.float64)
        self.numsum += 1
        self.last = v

    def averages(self) -> dict:
        """averages a<key> and spreads d<key> as runavs forms them (:248-263)."""
        rt = 1.0 / self.numsum
        out = {}
        for k in self.KEYS:
            a = rt * self.s[k]
            out["a" + k] = a
            out["d" + k] = math.sqrt(abs(rt * self.sq[k] - a * a))
        out["b0_mean"] = self.sb0 * rt
        return out


def stress_gpa(b0, tp, tk, vol: float) -> np.ndarray:
    """sigma(i,j) = -sum_k b0(j,k) (tp(i,k) + tk(i,k)) / vol, in GPa (src/analysis_m.F90:294-303)."""
    b0, tp, tk = (np.asarray(t, dtype=np.float64) for t in (b0, tp, tk))
    sig = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            sig[i, j] = -sum(b0[j, k] * (tp[i, k] + tk[i, k]) for k in range(3)) / vol
    return sig * EV_A3_TO_GPA


def runavs_text(step: int, sums: ThermoSums, temp: float, nm: int, bk: float, b0, tp, tk) -> str:
    """The running-averages block of runavs (src/analysis_m.F90:237-354) for an MD run."""
    a = sums.averages()
    L = sums.last
    atemp = a["ake"] / (1.5 * nm * bk)
    dtemp = a["dke"] / (1.5 * nm * bk)
    E = fortran_e
    out = [f" RUNNING AVERAGES AT STEP NO.{step:6d}", f" AVERAGED OVER {sums.numsum:6d} STEPS", "", "", ""]

    def line(lab, alab, x, ax, dx, unit, w1=16, d1=8, col=25):
        head = (lab + E(x, w1, d1)).ljust(col - 1)
        return head + alab + E(ax, 15, 6) + " +/-" + E(dx, 13, 4) + unit

    out.append(line(" PE:", "APE:", L["pe"], a["ape"], a["dpe"], " eV"))
    out.append(line(" KE:", "AKE:", L["ke"], a["ake"], a["dke"], " eV"))
    out.append(line(" TE:", "ATE:", L["te"], a["ate"], a["dte"], " eV"))
    out.append(("  T:" + fortran_f(temp, 11, 4)).ljust(24) + " AT:" + fortran_f(atemp, 11, 4) + " " * 5 + "+/-" +
               fortran_f(dtemp, 9, 4) + " K")
    out.append(line("  H:", " AH:", L["h"], a["ah"], a["dh"], " eV", 15, 6))
    out.append(line(" TH:", "ATH:", L["th"], a["ath"], a["dth"], " eV", 15, 6))
    out.append((" VOL:" + E(L["vol"], 15, 6)).ljust(23) + "AVOL:" + E(a["avol"], 15, 6) + " +/-" +
               E(a["dvol"], 13, 4) + " Ang**3")
    out.append("")
    sig = stress_gpa(b0, tp, tk, L["vol"])
    for i in range(3):
        out.append(f"TP+TK STRESS I={i + 1:2d}" + "".join(E(sig[i, j], 14, 6) for j in range(3)) + " GPa")
    out.append(f" VOLUME: {E(L['vol'], 14, 7).replace('E', 'D')} MD-BOX VECTORS (Ang) :")
    out.append("")
    b0 = np.asarray(b0, dtype=np.float64)
    for i in range(3):
        out.append(f"I={i + 1:3d}".ljust(6) + "MEAN VALUE:".ljust(20) + "".join(E(a["b0_mean"][j, i], 18, 10)
                                                                           for j in range(3)))
        out.append(" " * 6 + "INSTANTANEOUS VALUE:" + "".join(E(b0[j, i], 18, 10) for j in range(3)))
    return "\n".join(out) + "\n"
