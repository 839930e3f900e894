"""Generates the golden fixtures under tests/golden/ from the REFERENCE.

Runs oracle/_ref/ref_driver (the reference's own Fortran modules, built from
/root/reference/src by oracle/build_ref.sh) on seeded synthetic inputs and
stores inputs + outputs as small .npz fixtures (no pickles).  Run in the
build container only (it needs /root/reference for pot_026_026.in and the
built driver):

    python tests/golden/gen_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from moldy_amd import lattice  # noqa: E402

REF_POT = "/root/reference/src/potentials/pot_026_026.in"
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
OUT = os.path.dirname(os.path.abspath(__file__))


def r_grid() -> np.ndarray:
    base = np.linspace(0.05, 6.5, 1291)
    c1, c2 = float(np.float32(0.9)), float(np.float32(1.9))
    knots = [3.38247, 3.296475, 3.09582, 2.837835, 2.665845, 2.4824606625, 3.72645, 3.4398]
    special = [c1, c2, 1.8202052203699783, 6.0, 5.999999]
    for k in knots + [c1, c2]:
        special += [np.nextafter(k, 0.0), np.nextafter(k, 10.0), k - 1e-9, k + 1e-9]
    return np.concatenate([base, np.array(special)])


def rho_grid() -> np.ndarray:
    return np.concatenate([np.linspace(0.0, 60.0, 601), [1e-41, 1e-40, -1.0, 9999.0, 10000.0, 12000.0]])


class Reader:
    def __init__(self, buf):
        self.b = buf
        self.o = 0

    def f8(self, k):
        v = np.frombuffer(self.b, dtype="<f8", count=k, offset=self.o).copy()
        self.o += 8 * k
        return v

    def i4(self, k):
        v = np.frombuffer(self.b, dtype="<i4", count=k, offset=self.o).copy()
        self.o += 4 * k
        return v


def run_reference(sysdef, ivol=1, ntcm=10, nsteps=10, uselookup=0, nnbrs=150, rpad=0.3,
                  press=0.0, boxmass=50.0, flags=0, driver=DRIVER, pot_file=REF_POT):
    import refrun
    rg, rhog = r_grid(), rho_grid()
    n = sysdef.n
    _, buf = refrun.run(sysdef, ivol=ivol, ntcm=ntcm, nsteps=nsteps, uselookup=uselookup,
                        nnbrs=nnbrs, rpad=rpad, press=press, boxmass=boxmass, flags=flags,
                        rgrid=rg, rhogrid=rhog, driver=driver, pot_file=pot_file)
    R = Reader(buf)
    d = {}
    d["nlc"] = R.i4(3)
    d["rnear"], d["deltat"], d["bdel2"], d["press"], d["vol0"], d["pe0"] = R.f8(6)
    arr = R.f8(9 * n).reshape(9, n)
    for k, nm in enumerate(["rho0", "afrho0", "dafrho0", "fx0", "fy0", "fz0", "x20", "y20", "z20"]):
        d[nm] = arr[k]
    ten = R.f8(8 * 9).reshape(8, 9)
    for k, nm in enumerate(["tp0", "tk0", "tc0", "fb0", "tg0", "tgdot0", "tgid0", "b20"]):
        d[nm] = ten[k].reshape(3, 3, order="F")
    numn = R.i4(n)
    pi, pj = [], []
    for i in range(n):
        js = R.i4(int(numn[i]))
        pi.append(np.full(len(js), i))
        pj.append(js - 1)
    d["numn0"] = numn
    d["pairs0"] = np.stack([np.concatenate(pi), np.concatenate(pj)], axis=1)
    ng, nr = len(rg), len(rhog)
    d["rgrid"], d["rhogrid"] = rg, rhog
    for nm in ("vee", "dvee", "phi", "dphi"):
        d[nm] = R.f8(ng)
    for nm in ("emb", "demb"):
        d[nm] = R.f8(nr)
    d["trace"] = R.f8(6 * max(nsteps, 1)).reshape(max(nsteps, 1), 6)[:nsteps]
    fin = R.f8(14 * n).reshape(14, n)
    for k, nm in enumerate(["x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2", "fx", "fy",
                            "fz", "rho", "dafrho"]):
        d[nm + "_f"] = fin[k]
    ten = R.f8(7 * 9).reshape(7, 9)
    for k, nm in enumerate(["b0", "b1", "b2", "tp", "fb", "tgid", "tk"]):
        d[nm + "_f"] = ten[k].reshape(3, 3, order="F")
    return d


# fixture cases: name -> (system kwargs, run kwargs, keys to keep)
CASES = {
    # configs[0] shape: 10^3 bcc cells, 2000 atoms, 300 K, NVE, dt 1 fs, exact mode
    "c0_nve_2000": (dict(nc=10, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=1234),
                    dict(ivol=1, ntcm=10, nsteps=12, uselookup=0)),
    # Parrinello-Rahman NPT at 0 GPa on a sheared (triclinic) 8^3 box
    "npt_tric_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=99,
                           strain=np.array([[0.01, 0.004, -0.003], [0.004, -0.006, 0.002],
                                            [-0.003, 0.002, 0.003]])),
                      dict(ivol=0, ntcm=10, nsteps=12, uselookup=0, press=0.0, boxmass=50.0)),
    # lookup tables (set_up before potential_set_lookup)
    "lookup_2000": (dict(nc=10, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=77),
                    dict(ivol=1, ntcm=10, nsteps=3, uselookup=1)),
    # hot dense liquid-like state: close approaches reach the Biersack/join region
    "hot_1024": (dict(nc=8, a=2.98118, sigma=0.30, temp=2500.0, deltat_fs=0.5, seed=5),
                 dict(ivol=1, ntcm=5, nsteps=10, uselookup=0)),
}

KEEP_FIRST = ["rho0", "afrho0", "dafrho0", "fx0", "fy0", "fz0", "x20", "y20", "z20"]
KEEP_FINAL = ["x0_f", "y0_f", "z0_f", "x1_f", "y1_f", "z1_f", "fx_f", "fy_f", "fz_f"]
KEEP_TENS = ["tp0", "tk0", "tc0", "fb0", "tgid0", "b20", "b0_f", "b1_f", "b2_f", "tp_f", "fb_f",
             "tgid_f", "tk_f"]


def main():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if not os.path.exists(DRIVER):
        raise SystemExit("build the reference driver first: oracle/build_ref.sh")
    for name, (skw, rkw) in CASES.items():
        sd = lattice.make_system(**skw)
        d = run_reference(sd, **rkw)
        fx = {"b0": sd.b0, "s": sd.s, "x1": sd.x1, "mass": sd.mass, "z": sd.z,
              "deltat_fs": np.float64(sd.deltat_fs)}
        for k, v in rkw.items():
            fx["run_" + k] = np.asarray(v)
        for k in KEEP_FIRST + KEEP_FINAL + KEEP_TENS:
            fx[k] = d[k]
        for k in ("nlc", "rnear", "deltat", "bdel2", "press", "vol0", "pe0", "trace", "numn0"):
            fx[k] = np.asarray(d[k])
        fx["numn0"] = d["numn0"].astype(np.int16)
        fx["npairs0"] = np.int64(len(d["pairs0"]))
        fx["pairs0_sha256"] = np.array(po.pairset_digest(d["pairs0"]))
        if name in ("c0_nve_2000", "lookup_2000"):
            for k in ("rgrid", "rhogrid", "vee", "dvee", "phi", "dphi", "emb", "demb"):
                fx[k] = d[k]
        path = os.path.join(OUT, name + ".npz")
        np.savez_compressed(path, **fx)
        print(f"{name}: n={sd.n} pairs={len(d['pairs0'])} -> {path} "
              f"({os.path.getsize(path) // 1024} KiB)")


if __name__ == "__main__":
    main()
