"""Seeded synthetic inputs for the Fe workloads (bcc/fcc supercells).

Positions follow the replication order of the reference's system.in reader
(src/io_m.F90:226-248: loops over cells x, y, z, then basis atoms), are
fractional and never wrapped, and are perturbed by Gaussian Cartesian
displacements.  Velocities are Maxwell-Boltzmann, with the centre-of-mass
momentum removed and rescaled to exactly 3/2 N k T, and are stored the way
the reference stores them: X1 = (fractional velocity) * deltat
(src/particles_m.F90:58-60, src/moldin.F90:294-300).

Units are the reference's model units (src/constants_m.F90:84-113):
energy eV, length Angstrom, mass amu, time 10.1805 fs.
"""
from __future__ import annotations

import dataclasses

import numpy as np

FE_MASS = 55.845            # amu, src/particles_m.F90:152
FE_Z = 26
A_FE = 2.8665               # bcc lattice parameter (Angstrom)
ELECTRON = 1.602176487e-19
BOLTZMANN = 1.3806504e-23
BK = BOLTZMANN / ELECTRON   # eV per kelvin
TIME_TO_SEC = 1.018050530893617e-014
FS_TO_TIME = 1.0 / (TIME_TO_SEC / 1.0e-15)

BCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])
FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])


@dataclasses.dataclass
class System:
    """One periodic cell: b0 (3x3, columns are cell vectors), fractional
    positions s (N,3), X1 = sdot*dt (N,3), masses (N,), atomic numbers (N,)."""
    b0: np.ndarray
    s: np.ndarray
    x1: np.ndarray
    mass: np.ndarray
    z: np.ndarray
    deltat_fs: float
    info: dict | None = None  # how a generated cell was made (make_batch: nc, fcc, vacancy)

    @property
    def n(self) -> int:
        return int(self.s.shape[0])


def supercell(nc, a: float, basis=BCC_BASIS) -> tuple[np.ndarray, np.ndarray]:
    """Perfect supercell: returns (b0, s) with fractional positions in [0,1)."""
    nc = np.broadcast_to(np.asarray(nc, dtype=np.int64), (3,))
    b0 = np.diag(a * nc.astype(np.float64))
    ii, jj, kk = np.meshgrid(np.arange(nc[0]), np.arange(nc[1]), np.arange(nc[2]), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    s = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) / nc.astype(np.float64)
    return b0, s


def maxwell_boltzmann(rng: np.random.Generator, b0: np.ndarray, mass: np.ndarray,
                      temp: float, deltat_fs: float) -> np.ndarray:
    """Cartesian MB velocities at temp (K), COM removed, rescaled to 3/2 N k T;
    returned as X1 = B0^-1 v * deltat (fractional displacement per step)."""
    n = mass.shape[0]
    if temp <= 0.0 or n == 0:
        return np.zeros((n, 3))
    sig = np.sqrt(BK * temp / mass)[:, None]
    v = rng.standard_normal((n, 3)) * sig
    p = (v * mass[:, None]).sum(axis=0)
    v -= p / mass.sum()
    ke = 0.5 * (mass[:, None] * v * v).sum()
    v *= np.sqrt(1.5 * n * BK * temp / ke)
    dt = deltat_fs * FS_TO_TIME
    return (np.linalg.solve(b0, v.T).T) * dt


def make_system(nc, a: float = A_FE, basis=BCC_BASIS, sigma: float = 0.05,
                temp: float = 300.0, deltat_fs: float = 1.0, seed: int = 0,
                strain: np.ndarray | None = None, vacancy_frac: float = 0.0) -> System:
    """Seeded Fe supercell with Gaussian displacements sigma (Angstrom per
    Cartesian component), optional symmetric strain (b0 <- (I+e) b0) and a
    random fraction of removed atoms (vacancies)."""
    rng = np.random.default_rng(seed)
    b0, s = supercell(nc, a, basis)
    if strain is not None:
        b0 = (np.eye(3) + strain) @ b0
    if vacancy_frac > 0.0:
        nrem = int(round(vacancy_frac * s.shape[0]))
        if nrem > 0:
            keep = np.sort(rng.permutation(s.shape[0])[nrem:])
            s = s[keep]
    n = s.shape[0]
    if sigma > 0.0:
        dr = rng.standard_normal((n, 3)) * sigma
        s = s + np.linalg.solve(b0, dr.T).T
    mass = np.full(n, FE_MASS)
    x1 = maxwell_boltzmann(rng, b0, mass, temp, deltat_fs)
    return System(b0=b0, s=np.ascontiguousarray(s), x1=np.ascontiguousarray(x1),
                  mass=mass, z=np.full(n, FE_Z, dtype=np.int32), deltat_fs=deltat_fs)


def liquid_lattice_parameter(density_g_cm3: float) -> float:
    """bcc lattice parameter giving the mass density (2 atoms per cell)."""
    avogadro = 6.02214179e23
    vol_cm3 = 2.0 * FE_MASS / avogadro / density_g_cm3
    return (vol_cm3 * 1e24) ** (1.0 / 3.0)


def make_batch(ncell: int = 4096, seed: int = 7, nmin: int = 3, nmax: int = 10,
               sigma: float = 0.1, max_strain: float = 0.03, max_vacancy: float = 0.02,
               a_bcc: float = A_FE) -> list[System]:
    """Independent periodic Fe cells of uneven size: bcc or fcc (probability 1/2)
    n^3 supercells with n uniform in [nmin, nmax], fcc at the bcc volume per atom
    (a_fcc = a_bcc 2^(1/3)), a random symmetric strain with each of the six
    components uniform in [-max_strain, max_strain], a random vacancy fraction
    uniform in [0, max_vacancy] and Gaussian displacements sigma (Angstrom)."""
    rng = np.random.default_rng(seed)
    cells = []
    for c in range(ncell):
        n = int(rng.integers(nmin, nmax + 1))
        fcc = bool(rng.random() < 0.5)
        e = rng.uniform(-max_strain, max_strain, 6)
        strain = np.array([[e[0], e[3], e[4]], [e[3], e[1], e[5]], [e[4], e[5], e[2]]])
        vac = float(rng.uniform(0.0, max_vacancy))
        cell = make_system(nc=n, a=a_bcc * 2 ** (1 / 3) if fcc else a_bcc,
                           basis=FCC_BASIS if fcc else BCC_BASIS, sigma=sigma, temp=0.0,
                           seed=int(rng.integers(1 << 31)), strain=strain, vacancy_frac=vac)
        cell.info = {"nc": n, "fcc": fcc, "vacancy": vac}
        cells.append(cell)
    return cells
