"""Initial state of a new MOLDY run, as moldin builds it on the host.

* Ranmar: the reference's random number generator (src/random_m.F90:73-143),
  Marsaglia-Zaman-Tsang, started by rstart(97,33,21,82) (src/moldin.F90:175).
* read_system: system.in -- basis atoms replicated nt(1) x nt(2) x nt(3)
  times in the reference's atom order, each coordinate displaced by
  (rand1()-1/2) dsp / B0(k,k) (src/io_m.F90:73-256); three draws per atom
  even when dsp = 0.
* setvel: random velocities, centre-of-mass drift removed, scaled to the
  required kinetic energy (src/thermostat_m.F90:191-306; the atom part, used
  for ivol < 1 and ivol == 4, where the routine returns before the box part).

Host code: O(N) once per run, off the device path.
"""
from __future__ import annotations

import numpy as np

BK = 1.3806504e-23 / 1.602176487e-19   # src/constants_m.F90:93


class Ranmar:
    def __init__(self, i: int = 97, j: int = 33, k: int = 21, l: int = 82):
        self.u = [0.0] * 97
        for ii in range(97):
            s, t = 0.0, 0.5
            for _ in range(24):
                m = ((i * j) % 179) * k % 179
                i, j, k = j, k, m
                l = (53 * l + 1) % 169
                if (l * m) % 64 >= 32:
                    s = s + t
                t = 0.5 * t
            self.u[ii] = s
        # REAL(4) quotients of 24-bit integers, exact in binary
        self.c = 362436.0 / 16777216.0
        self.cd = 7654321.0 / 16777216.0
        self.cm = 16777213.0 / 16777216.0
        self.iu, self.ju = 97, 33

    def rand1(self) -> float:
        u = self.u
        uni = u[self.iu - 1] - u[self.ju - 1]
        if uni < 0.0:
            uni += 1.0
        u[self.iu - 1] = uni
        self.iu -= 1
        if self.iu == 0:
            self.iu = 97
        self.ju -= 1
        if self.ju == 0:
            self.ju = 97
        self.c -= self.cd
        if self.c < 0.0:
            self.c += self.cm
        uni -= self.c
        if uni < 0.0:
            uni += 1.0
        return uni


def parse_system(text: str):
    """system.in: nbasis; nt(3); the unit volume's three vectors b0(:,i); then
    per basis atom x y z (fractional in the unit volume), atomic number, mass."""
    toks = [ln.split("#")[0].split() for ln in text.splitlines()]
    toks = [t for t in toks if t]
    nbasis = int(toks[0][0])
    nt = [int(v) for v in toks[1][:3]]
    cell = np.array([[float(v.replace("d", "e").replace("D", "E")) for v in toks[2 + i][:3]] for i in range(3)]).T
    basis = []
    for b in range(nbasis):
        t = toks[5 + b]
        basis.append((float(t[0]), float(t[1]), float(t[2]), int(t[3]), float(t[4])))
    return nbasis, nt, cell, basis


def read_system(text: str, rng: Ranmar, dsp: float = 0.0):
    """Whole-box B0, fractional positions (n,3), atomic numbers, masses --
    the replication loop of read_system (src/io_m.F90:216-245)."""
    nbasis, nt, cell, basis = parse_system(text)
    b0 = cell * np.array(nt, dtype=np.float64)[None, :]        # b0(:,i) = b0(:,i) * nt(i)
    rt = [1.0 / nt[0], 1.0 / nt[1], 1.0 / nt[2]]
    n = nbasis * nt[0] * nt[1] * nt[2]
    s = np.empty((n, 3))
    z = np.empty(n, dtype=np.int32)
    mass = np.empty(n)
    l = 0
    for i in range(nt[0]):
        for j in range(nt[1]):
            for k in range(nt[2]):
                for (xb, yb, zb, zn, mb) in basis:
                    s[l, 0] = (rng.rand1() - 0.5) * dsp / b0[0, 0] + rt[0] * (i + xb)
                    s[l, 1] = (rng.rand1() - 0.5) * dsp / b0[1, 1] + rt[1] * (j + yb)
                    s[l, 2] = (rng.rand1() - 0.5) * dsp / b0[2, 2] + rt[2] * (k + zb)
                    z[l] = zn
                    mass[l] = mb
                    l += 1
    return b0, s, z, mass


def setvel(rng: Ranmar, z, mass, b0, deltat: float, rqke: float) -> np.ndarray:
    """x1 = v dt / L of setvel's atom part (src/thermostat_m.F90:191-255)."""
    n = len(z)
    x1 = np.zeros((n, 3))
    moving = (np.asarray(z) != 0) & (np.asarray(mass) >= float(np.float32(0.1)))
    for i in np.flatnonzero(moving):
        x1[i, 0] = rng.rand1() - 0.5
        x1[i, 1] = rng.rand1() - 0.5
        x1[i, 2] = rng.rand1() - 0.5
    nmove = int(moving.sum())
    sx = sy = sz = 0.0
    for i in range(n):                       # sequential sums, all atoms
        sx = sx + x1[i, 0] * mass[i]
        sy = sy + x1[i, 1] * mass[i]
        sz = sz + x1[i, 2] * mass[i]
    sx, sy, sz = sx / nmove, sy / nmove, sz / nmove
    b = np.asarray(b0, dtype=np.float64)
    s3 = 0.0
    for i in np.flatnonzero(moving):
        x1[i, 0] = x1[i, 0] - sx / mass[i]
        x1[i, 1] = x1[i, 1] - sy / mass[i]
        x1[i, 2] = x1[i, 2] - sz / mass[i]
        a, c, d = x1[i]
        rx = b[0, 0] * a + b[0, 1] * c + b[0, 2] * d
        ry = b[1, 0] * a + b[1, 1] * c + b[1, 2] * d
        rz = b[2, 0] * a + b[2, 1] * c + b[2, 2] * d
        s3 = s3 + mass[i] * (rx * rx + ry * ry + rz * rz)
    ke = s3 / (deltat * deltat) / 2.0
    # float(nmove)/float(nm) and its sqrt are REAL(4)
    scale = np.sqrt(rqke / ke) * float(np.sqrt(np.float32(nmove) / np.float32(n)))
    return x1 * scale


def setvel_box(rng: Ranmar, ivol: int, nm: int, temprq: float, bdel2: float) -> np.ndarray:
    """b1 of setvel's box part (src/thermostat_m.F90:256-299), drawn after the
    atom part from the same stream.  Only ivol 1, 2, 3 and 5 reach it (the box
    of those runs does not move, so b1 enters the box kinetic energy and the
    checkpoint); otherwise b1 is zero and nothing is drawn.  Rows are drawn in
    order, each row loses its sum / nm, and the whole is scaled to the box
    temperature temprq, boxtem = (sum b1^2 / bdel2) / (4.5 bk)."""
    b1 = np.zeros((3, 3))
    if ivol < 1 or ivol == 4:
        return b1
    s3 = 0.0
    for i in range(3):
        s = 0.0
        for j in range(3):
            b1[i, j] = rng.rand1() - 0.5
            s = s + b1[i, j]
        s = s / nm
        for j in range(3):
            b1[i, j] = b1[i, j] - s
            s3 = s3 + b1[i, j] * b1[i, j]
    boxtem = (s3 / bdel2) / (4.5 * BK)
    return b1 * np.sqrt(temprq / boxtem)


def rqke_init(temprq: float, nm: int) -> float:
    """RQKE of init_thermostat_m (src/thermostat_m.F90:97)."""
    return 1.5 * BK * temprq * (nm - 1)
