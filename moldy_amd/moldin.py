"""MOLDY's main program for molecular dynamics, running the hot path on the GPU.

    python -m moldy_amd.moldin [--material fetransmetal|iron_carbon] [--dir DIR]

Follows src/moldin.F90 for MD runs (iquen = 0), in its order:
  rstart(97,33,21,82); read_params (params.in); read_system (file_system);
  potential (lookup or exact) and link-cell sizes; restart 0 (setvel and its
  box part, scalev) or -1/1 (read the checkpoint); strain loops: zero the
  sums, strain the box (ivol != 0), adjust press/temprq, first force and
  accelerations; the MD loop
  (velocity Verlet, or PREDIC/FORCE/CORREC for iverlet = 0; update_therm_avs;
  relink), with the thermodynamic sums sampled nprint/10 times between
  printouts and runavs every nprint steps; checkpoints every nchkpt steps and
  at the end; energy_calc and system.out.
The textout file receives the running-averages blocks and the final box.

The material is a run-time choice here (the reference compiles one in):
fetransmetal reads pot_026_026.in from the run directory (its Fe-Fe knots),
iron_carbon reads iron.in when present.

Checkpoint files keep write_checkpoint_file's record layout (Fortran
sequential records with 4-byte length markers; src/moldin.F90:1053-1070):
parameters, (x0,y0,z0,b0), (x1,y1,z1,b1), (x2,y2,z2,b2), (x3,y3,z3,b3),
atomic numbers, masses, species indices, species list, thermodynamic sums.
The reference's first record is a memory image of its simparameters type,
which depends on the compiler; this program writes its own parameter record
(MOLDYAMD tag, counters and scalars, little-endian) and, reading a
checkpoint, uses the counters of that record when it is one of its own and
those of params.in otherwise.  All other records are the reference's.

The Nose thermostat (nose > 0, or therm_con > 0 converted to a Nose time)
runs scalev after setvel and whenever the accumulated time reaches one fs.

iquen = 1 relaxes the structure with moldy_amd.quench (conjugate gradients
on the enthalpy) and writes the FINAL U line and system.out; iquen = 2 runs
MD with the box quench (box velocities that oppose their acceleration zeroed
after every step).

New runs draw setvel's box velocities too (ivol 1, 2, 3, 5, whose box does
not move: they enter the box kinetic energy and the checkpoint).  Each of the
strainloops loops strains the box by strx (ivol != 0), steps press and temprq
by pressstep and tempsp, and runs nsteps steps (or one quench).

Not supported (errors): time-step adaptation (iverlet = 2); metropolis, PKA
and the dump files are ignored.  Steps run one md_steps call each, so that
the per-step f^2 sums of the checkpoint are kept.
"""
from __future__ import annotations

import argparse
import os
import struct
import sys

import numpy as np

from . import engine, initial, io, lattice
from .params import read_params

_PTAG = b"MOLDYAMD"


def _rec(fh, data: bytes) -> None:
    fh.write(struct.pack("<i", len(data)))
    fh.write(data)
    fh.write(struct.pack("<i", len(data)))


def write_checkpoint(path: str, counters: dict, scalars: dict, s, x1, x2, x3, b, z, mass, aidx, ispec,
                     thm: np.ndarray, numsum: int) -> None:
    """s, x1, x2, x3: (n,3); b: the four box matrices b0..b3; thm: the 24
    doubles SPE..SVOLPE, SB0 of the thermodynamic sums."""
    keys_i = ("nlcx", "nlcy", "nlcz", "prevsteps", "currentstep", "laststep", "lastprint", "lastchkpt", "ntc",
              "iverlet", "ivol")
    keys_d = ("deltat", "press", "temprq", "bmass", "bdel2")
    head = _PTAG + struct.pack("<11i", *[int(counters[k]) for k in keys_i]) + \
        struct.pack("<5d", *[float(scalars[k]) for k in keys_d])
    with open(path, "wb") as fh:
        _rec(fh, head)
        for arr, bm in ((s, b[0]), (x1, b[1]), (x2, b[2]), (x3, b[3])):
            a = np.asarray(arr, dtype="<f8")
            _rec(fh, a[:, 0].tobytes() + a[:, 1].tobytes() + a[:, 2].tobytes() +
                 np.asarray(bm, dtype="<f8").ravel(order="F").tobytes())
        _rec(fh, np.asarray(z, dtype="<i4").tobytes())
        _rec(fh, np.asarray(mass, dtype="<f8").tobytes())
        _rec(fh, np.asarray(aidx, dtype="<i4").tobytes())
        _rec(fh, np.asarray(ispec, dtype="<i4").tobytes())
        _rec(fh, np.asarray(thm, dtype="<f8").tobytes() + struct.pack("<i", int(numsum)))


def read_checkpoint(path: str, n: int) -> dict:
    """Records of a checkpoint (this program's or the reference's)."""
    raw = open(path, "rb").read()
    recs, o = [], 0
    while o < len(raw):
        ln = struct.unpack_from("<i", raw, o)[0]
        recs.append(raw[o + 4:o + 4 + ln])
        o += ln + 8
    if len(recs) < 10:
        raise ValueError("checkpoint: expected 10 records, found %d" % len(recs))
    out = {}
    if recs[0][:8] == _PTAG:
        vi = struct.unpack_from("<11i", recs[0], 8)
        vd = struct.unpack_from("<5d", recs[0], 8 + 44)
        out["counters"] = dict(zip(("nlcx", "nlcy", "nlcz", "prevsteps", "currentstep", "laststep", "lastprint",
                                    "lastchkpt", "ntc", "iverlet", "ivol"), vi))
        out["scalars"] = dict(zip(("deltat", "press", "temprq", "bmass", "bdel2"), vd))
    for k, nm in enumerate(("0", "1", "2", "3")):
        v = np.frombuffer(recs[1 + k], "<f8")
        if v.size != 3 * n + 9:
            raise ValueError("checkpoint does not hold %d atoms" % n)
        out["x" + nm] = v[:3 * n].reshape(3, n).T.copy()
        out["b" + nm] = v[3 * n:].reshape(3, 3, order="F").copy()
    out["z"] = np.frombuffer(recs[5], "<i4").copy()
    out["mass"] = np.frombuffer(recs[6], "<f8").copy()
    out["aidx"] = np.frombuffer(recs[7], "<i4").copy()
    out["ispec"] = np.frombuffer(recs[8], "<i4").copy()
    out["thm"] = np.frombuffer(recs[9][:192], "<f8").copy()
    out["numsum"] = struct.unpack_from("<i", recs[9], 192)[0]
    return out


class Sums(io.ThermoSums):
    """Thermodynamic sums with the reference's per-step f^2 sums (moldin.F90:762-765)
    and the volume-energy product (analysis_m.F90:155)."""

    def __init__(self):
        super().__init__()
        self.sf2 = self.sf2sq = self.svolpe = 0.0

    def update(self, pe, ke, te, h, th, vol, b0):
        super().update(pe, ke, te, h, th, vol, b0)
        self.svolpe = self.svolpe + vol * pe

    def as_record(self) -> np.ndarray:
        s, q = self.s, self.sq
        return np.concatenate([[s["pe"], s["ke"], s["te"], s["h"], s["th"], self.sf2, s["vol"], q["pe"], q["ke"],
                                q["te"], q["h"], q["th"], self.sf2sq, q["vol"], self.svolpe],
                               np.asarray(self.sb0).ravel(order="F")])


def run(workdir: str = ".", material: str = "fetransmetal") -> dict:
    """Runs the MD job described by workdir/params.in; returns a summary."""
    p = read_params(os.path.join(workdir, "params.in"))
    if p.iverlet not in (0, 1):
        raise NotImplementedError("iverlet must be 0 or 1")
    out = open(os.path.join(workdir, p.file_textout), "w")
    rng = initial.Ranmar(97, 33, 21, 82)                       # moldin.F90:175
    rqke = initial.rqke_init(p.temprq, p.nm)                    # init_thermostat_m
    text = open(os.path.join(workdir, p.file_system)).read()
    b0, s, z, mass = initial.read_system(text, rng, p.dsp)
    n = len(z)
    if p.nm and p.nm != n:
        raise ValueError(f"params.in nm = {p.nm} but {p.file_system} holds {n} atoms")
    # species bookkeeping of read_system for the checkpoint (one index per distinct Z)
    ispec = []
    for zz in z:
        if int(zz) not in ispec:
            ispec.append(int(zz))
    aidx = np.array([ispec.index(int(zz)) + 1 for zz in z], dtype=np.int32)
    out.write(f"  STRUCTURE READ IN {len(initial.parse_system(text)[3]):12d}  BASIS ATOMS;"
              f"{n // len(initial.parse_system(text)[3]):12d}  CELLS;{n:12d}   ATOMS.\n")
    out.write(f" {p.title1}\n {p.title2}\n")
    x1 = np.zeros((n, 3))
    x2 = np.zeros((n, 3))
    x3 = np.zeros((n, 3))
    b = [np.asarray(b0, dtype=np.float64), np.zeros((3, 3)), np.zeros((3, 3)), np.zeros((3, 3))]
    counters = dict(prevsteps=0, currentstep=0, laststep=p.nsteps, lastprint=0, lastchkpt=0, ntc=p.ntc)
    if p.restart == 0:
        if p.iquen != 1:                                        # moldin.F90:307
            x1 = initial.setvel(rng, z, mass, b[0], p.deltat, rqke)
            b[1] = initial.setvel_box(rng, p.ivol, n, p.temprq, p.bdel2)
        out.write("\n          EQUILIBRATION STARTING FROM REGULAR LATTICE\n\n\n")
    else:
        ck = read_checkpoint(os.path.join(workdir, p.file_checkpointread), n)
        s = ck["x0"]
        b = [ck["b0"], ck["b1"], ck["b2"], ck["b3"]]
        if p.restart == 1:
            x1, x2, x3 = ck["x1"], ck["x2"], ck["x3"]
        else:   # restart -1: positions only (moldin.F90:343-352)
            b = [b[0], np.zeros((3, 3)), np.zeros((3, 3)), np.zeros((3, 3))]
            if p.iquen != 1:                                    # moldin.F90:338
                x1 = initial.setvel(rng, z, mass, b[0], p.deltat, rqke)
                b[1] = initial.setvel_box(rng, p.ivol, n, p.temprq, p.bdel2)
        if "counters" in ck:
            c = ck["counters"]
            counters.update(prevsteps=c["currentstep"], currentstep=c["currentstep"], lastprint=c["lastprint"],
                            lastchkpt=c["lastchkpt"], ntc=c["ntc"])
            counters["laststep"] = counters["prevsteps"] + p.nsteps
    nose = p.nose
    if nose != 0 and p.therm_con > 0:   # thermal conductivity -> Nose time (moldin.F90:206-215)
        bd = b[0][0, 0] * 1e-10 + b[0][1, 1] * 1e-10 + b[0][2, 2] * 1e-10
        nose = (3.0 * n * 1.3806504e-23) / (2.0 * p.therm_con * bd) / 1.0e-15
    sd = lattice.System(b0=b[0], s=s, x1=x1, mass=mass, z=z, deltat_fs=p.deltat_fs)
    eng = engine.Engine(sd, ivol=p.ivol, rpad=p.rpad, nnbrs=p.nnbrs or 150, ntcm=max(p.ntcm, 1),
                        press_gpa=0.0, boxmass=p.bmass / max(n, 1), lookup=5000 if p.uselookup else 0,
                        material=material, iverlet=p.iverlet, rebuild=0,
                        iron_in=_iron_in(workdir) if material == "iron_carbon" else None,
                        coeffs=_pot_file(workdir) if material == "fetransmetal" else None)
    try:
        for nm, arr in (("x2", x2), ("x3", x3)):
            for k, ax in enumerate("xyz"):
                eng.upload(ax + nm[1], arr[:, k])
        for k, t in enumerate(("b1", "b2", "b3")):
            eng.set_tensor(t, b[k + 1])
        eng.ntc = counters["ntc"]
        if p.restart in (0, -1) and nose > 0 and p.iquen != 1:
            eng.scalev(nose, rqke)                       # after setvel (moldin.F90:308, 339)
        if p.restart == 0:
            # ke as setvel leaves it (RQKE), or as scalev recomputes it before
            # scaling (moldin.F90:309-310)
            ke0 = _scalev_ke(x1, b[0], mass, z, p.deltat) if (nose > 0 and p.iquen != 1) else rqke
            out.write(" " * 16 + "INITIAL TEMPERATURE".ljust(23) + io.fortran_f(ke0 / (1.5 * initial.BK * n), 14, 8)
                      + " K\n\n\n")
        eng.linkup()
        out.write("\n" + " " * 10 + " SET UP LINK CELLS\n\n\n")
        fs_ref = float(np.float32(0.098227))    # one fs in model units, a REAL(4) literal (moldin.F90:132)
        thermotimer = 0.0
        every = p.nprint // 10 if (p.nprint >= 10 and p.nprint <= p.nsteps) else 0
        summary = {}
        press, temprq = p.press, p.temprq
        for _loop in range(max(p.strainloops, 1)):
            # one strain loop (moldin.F90:418-477): sums zeroed, box strained
            # for ivol != 0, pressure and temperature stepped, first force and
            # accelerations, then nsteps MD steps
            out.write(f" Strainloop: T= {io.fortran_f(temprq, 12, 4)} dT= {io.fortran_f(p.tempsp, 12, 4)}\n")
            out.write(f" Strainloop: P= {io.fortran_e(press, 14, 6)} dP= {io.fortran_e(p.pressstep, 14, 6)}\n")
            sums = Sums()
            if p.ivol != 0:
                eng.set_tensor("b0", np.asarray(p.strx) @ eng.tensor("b0"))
                bb = eng.tensor("b0")
                out.write(" " * 10 + "TEMPRQ =".ljust(10) + io.fortran_f(temprq, 10, 2) + " " * 10 +
                          "b0*strx:".ljust(10) + "\n")
                for i in range(3):
                    out.write("".join(io.fortran_e(bb[k, i], 17, 8).replace("E", "D") for k in range(3)) + "\n")
                out.write("\n")
            press = press + p.pressstep
            temprq = temprq + p.tempsp
            rqke_md = 1.5 * initial.BK * temprq * (n - 3)       # moldin.F90:437
            eng.set_pressure(press)
            eng.md_init()
            if p.iquen == 1:   # static relaxation instead of MD (moldin.F90:566-572)
                from . import quench as q
                r = q.quench(eng, p.ivol, press)
                out.write(" QUENCH, FINDING MINIMUM ENERGY CONFIGURATION OF%5d ATOMS\n" % n)
                out.write("FINAL U= %s V=%s P=%s H=%s\n" % (
                    io.fortran_e(r["pe"] / n, 18, 10).replace("E", "D"), io.fortran_f(r["vol"] / n, 10, 6),
                    io.fortran_e(press / initial_gpa(), 12, 4).replace("E", "D"),
                    io.fortran_e(r["h"] / n, 13, 7).replace("E", "D")))
                summary = r
                continue
            for istep in range(counters["prevsteps"] + 1, counters["laststep"] + 1):
                counters["currentstep"] = istep
                counters["lastprint"] += 1
                counters["lastchkpt"] += 1
                pe, ke, te, h, th, vol = eng.md_steps(1)[0]
                sc = eng.scalars()
                sums.sf2 = sums.sf2 + sc["s1"]       # moldin.F90:762-765, every step
                sums.sf2sq = sums.sf2sq + sc["s2"]
                if every and counters["lastprint"] % every == 0:
                    sums.update(pe, ke, te, h, th, vol, eng.tensor("b0"))
                    if counters["lastprint"] % p.nprint == 0:
                        out.write("\n" + io.runavs_text(istep - counters["prevsteps"], sums,
                                                        ke / (1.5 * initial.BK * n), n, initial.BK,
                                                        eng.tensor("b0"), eng.tensor("tp"), eng.tensor("tk")))
                if p.nchkpt > 0 and counters["lastchkpt"] % p.nchkpt == 0:
                    _checkpoint(eng, p, counters, sums, z, mass, aidx, ispec, workdir)
                    write_efs(eng, workdir)                           # moldin.F90:815
                    out.write("Checkpoint at step:%6d\n\n\n\n" % istep)
                thermotimer = thermotimer + p.deltat                  # moldin.F90:828-835
                if thermotimer >= fs_ref:
                    thermotimer = thermotimer - fs_ref
                    if nose != 0:
                        eng.scalev(nose, rqke_md)
                if p.iquen == 2:
                    _box_quench(eng)
                summary = dict(pe=pe, ke=ke, te=te, h=h, th=th, vol=vol, steps=istep)
            counters["prevsteps"] = counters["currentstep"]             # moldin.F90:862, 902
            counters["laststep"] = counters["currentstep"] + p.nsteps
        if p.iquen == 1:
            en = eng.energy_calc()
            io.write_system_out(os.path.join(workdir, "system.out"), eng.tensor("b0"), eng.array("x0"),
                                eng.array("y0"), eng.array("z0"), z, mass, en)
            summary["en_atom"] = en
            return summary
        out.write("\n" + " " * 20 + "QUIT MOLECULAR DYNAMICS LOOP\n")
        en = eng.energy_calc()
        if p.nchkpt > 0:
            _checkpoint(eng, p, counters, sums, z, mass, aidx, ispec, workdir)
        bb = eng.tensor("b0")
        out.write(f" VOLUME: {io.fortran_e(eng.scalar('vol'), 14, 7).replace('E', 'D')} MD-BOX VECTORS (Ang) :\n\n")
        for i in range(3):
            out.write(f"I={i + 1:3d}".ljust(6) + "FINAL VALUE:".ljust(20) +
                      "".join(io.fortran_e(bb[j, i], 18, 10) for j in range(3)) + "\n\n")
        io.write_system_out(os.path.join(workdir, "system.out"), bb, eng.array("x0"), eng.array("y0"),
                            eng.array("z0"), z, mass, en)
        summary["en_atom"] = en
        return summary
    finally:
        out.close()
        eng.close()


def _scalev_ke(x1, b0, mass, z, deltat) -> float:
    """Kinetic energy scalev forms before it scales (src/thermostat_m.F90:336-363):
    the mean of x1 over all atoms removed from the moving ones."""
    moving = (np.asarray(z) != 0) & (np.asarray(mass) >= float(np.float32(0.1)))
    v = (np.asarray(x1)[moving] - np.asarray(x1).sum(axis=0) / float(len(z))) @ np.asarray(b0).T
    return float((np.asarray(mass)[moving] * (v * v).sum(axis=1)).sum()) / deltat ** 2 / 2.0


def _box_quench(eng) -> None:
    """iquen = 2 (src/moldin.F90:845-853): after each MD step every box
    velocity element b1(i,j) that is not moving along its acceleration
    b2(i,j) (b1 b2 <= 0) is set to zero, so the box relaxes while the atoms
    keep their dynamics."""
    b1, b2 = eng.tensor("b1"), eng.tensor("b2")
    eng.set_tensor("b1", np.where(b1 * b2 > 0.0, b1, 0.0))


def initial_gpa() -> float:
    from .params import GPA_TO_PRESS
    return GPA_TO_PRESS


def _checkpoint(eng, p, counters, sums, z, mass, aidx, ispec, workdir):
    path = os.path.join(workdir, p.file_checkpointwrite or "chkp_out.dat")
    arr = {nm: np.stack([eng.array(a + nm[1:]) for a in "xyz"], axis=1) for nm in ("x0", "x1", "x2", "x3")}
    write_checkpoint(path, dict(nlcx=eng.params.nlcx, nlcy=eng.params.nlcy, nlcz=eng.params.nlcz,
                                iverlet=p.iverlet, ivol=p.ivol, **counters),
                     dict(deltat=p.deltat, press=p.press, temprq=p.temprq, bmass=p.bmass, bdel2=p.bdel2),
                     arr["x0"], arr["x1"], arr["x2"], arr["x3"],
                     [eng.tensor(t) for t in ("b0", "b1", "b2", "b3")], z, mass, aidx,
                     np.array(ispec, dtype=np.int32), sums.as_record(), sums.numsum)


def write_efs(eng, workdir: str) -> None:
    """efs.out, as write_energy_forces_stress (src/dynamics_m.F90:588-619)
    writes it at each in-run checkpoint: the word energy and pe, the
    Cartesian force B0 f of every atom (reference order), then sigma(i,j) =
    -sum_k b0(j,k) tp(i,k) / vol in column order; whitespace-separated, as
    list-directed output is read."""
    b0, tp = eng.tensor("b0"), eng.tensor("tp")
    sigma = -(tp @ b0.T) / eng.scalar("vol")
    with open(os.path.join(workdir, "efs.out"), "w") as fh:
        fh.write(" energy %25.16E\n" % eng.scalar("pe"))
        for row in eng.forces_cart():
            fh.write(" %25.16E %25.16E %25.16E\n" % tuple(row))
        fh.write(" " + " ".join("%25.16E" % v for v in sigma.ravel(order="F")) + "\n")


def read_efs(path: str):
    """(energy, forces (n, 3), sigma (3, 3)) of an efs.out."""
    toks = open(path).read().split()
    assert toks[0] == "energy"
    vals = np.array([float(t.replace("D", "E").replace("d", "e")) for t in toks[1:]])
    n = (len(vals) - 10) // 3
    return vals[0], vals[1:1 + 3 * n].reshape(n, 3), vals[1 + 3 * n:].reshape(3, 3, order="F")


def _iron_in(workdir: str):
    path = os.path.join(workdir, "iron.in")
    if not os.path.exists(path):
        return None
    toks = open(path).read().split("\n", 1)[1].replace("d", "e").replace("D", "E").split()
    return np.array([float(t) for t in toks[:38]])


def _pot_file(workdir: str):
    from .potential import read_avtf
    path = os.path.join(workdir, "pot_026_026.in")
    return read_avtf(path) if os.path.exists(path) else None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--dir", default=".", help="run directory holding params.in")
    ap.add_argument("--material", default="fetransmetal", choices=sorted(engine.MATERIAL_RANGE))
    a = ap.parse_args(argv)
    run(a.dir, a.material)
    return 0


if __name__ == "__main__":
    sys.exit(main())
