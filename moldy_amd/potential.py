"""Fe-Fe potential coefficients and the pot_XXX_YYY.in reader.

The file layout is the reference's AVTF coefficient format
(src/potentials/pot_026_026.in:1-11, read by src/fetransmetal.F90:603-766):
a title line, two atomic numbers, the pair and density knot counts, the
radius range, then pair amplitudes, pair knots, density amplitudes and
density knots, one group per line, '#' starting a trailing comment.
"""
from __future__ import annotations

import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def read_avtf(path: str) -> dict:
    vals = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    body = [ln.split("#", 1)[0].strip() for ln in lines[1:]]
    body = [b for b in body if b]
    z1, z2 = int(body[0]), int(body[1])
    nv, np_ = int(body[2]), int(body[3])
    rmin, rmax = float(body[4]), float(body[5])
    av = np.array([float(v) for v in body[6].split()][:nv])
    rv = np.array([float(v) for v in body[7].split()][:nv])
    ap = np.array([float(v) for v in body[8].split()][:np_])
    rp = np.array([float(v) for v in body[9].split()][:np_])
    del vals
    return {"z1": z1, "z2": z2, "av": av, "rv": rv, "ap": ap, "rp": rp,
            "rmin": rmin, "rmax": rmax}


FE_POTENTIAL = read_avtf(os.path.join(DATA_DIR, "pot_026_026.in"))

# Range beyond which V, V', phi and phi' vanish identically (outermost knot).
FE_INTERACTION_RANGE = float(max(FE_POTENTIAL["rv"].max(), FE_POTENTIAL["rp"].max()))


def theta_vector(c: dict | None = None) -> np.ndarray:
    """16 knot parameters (6 a_k, 6 r_k, 2 A_k, 2 R_k) in that order."""
    c = FE_POTENTIAL if c is None else c
    return np.concatenate([c["av"], c["rv"], c["ap"], c["rp"]]).astype(np.float64)


def coeffs_from_theta(theta) -> dict:
    t = np.asarray(theta, dtype=np.float64)
    return {"av": t[0:6].copy(), "rv": t[6:12].copy(), "ap": t[12:14].copy(),
            "rp": t[14:16].copy(), "z1": 26, "z2": 26, "rmin": 0.0,
            "rmax": float(max(t[6:12].max(), t[14:16].max()))}
