"""Generates the golden fixtures under tests/golden/ from the REFERENCE.

Runs oracle/_ref/ref_driver and ref_driver_fec (the reference's own Fortran
modules with the fetransmetal or the iron_carbon material, built from
/root/reference/src by oracle/build_ref.sh) on seeded synthetic inputs and
stores inputs + outputs as small .npz fixtures (no pickles).  "bcc_iron"
extracts the numbers printed by the reference's own shipped test
(tests/reference/bcc_iron/output66.txt, system.out) with their inputs.  Run in the
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
DRIVER_FEC = os.path.join(ROOT, "oracle", "_ref", "ref_driver_fec")
OUT = os.path.dirname(os.path.abspath(__file__))
BCC_IRON = "/root/reference/tests/reference/bcc_iron"
REF_TESTS = "/root/reference/tests/reference"


def r_grid(material: str = "fetransmetal") -> np.ndarray:
    base = np.linspace(0.05, 6.5, 1291)
    if material == "iron_carbon":
        # Biersack/join boundary 1.0, join/knot boundary REAL(4) 2.05, knots, table end 5.4
        special = [1.0, float(np.float32(2.05)), 2.05, 5.4, 5.399999]
        knots = [5.3, 4.7, 4.2, 3.7, 3.3, 3.0, 2.8, 2.7, 2.6, 2.5, 2.4, 2.3, 2.2, 3.2]
        for k in knots + [1.0, float(np.float32(2.05))]:
            special += [np.nextafter(k, 0.0), np.nextafter(k, 10.0), k - 1e-9, k + 1e-9]
        return np.concatenate([base, np.array(special)])
    c1, c2 = float(np.float32(0.9)), float(np.float32(1.9))
    knots = [3.38247, 3.296475, 3.09582, 2.837835, 2.665845, 2.4824606625, 3.72645, 3.4398]
    special = [c1, c2, 1.8202052203699783, 6.0, 5.999999]
    for k in knots + [c1, c2]:
        special += [np.nextafter(k, 0.0), np.nextafter(k, 10.0), k - 1e-9, k + 1e-9]
    return np.concatenate([base, np.array(special)])


def rho_grid(material: str = "fetransmetal") -> np.ndarray:
    g = np.concatenate([np.linspace(0.0, 60.0, 601), [1e-41, 1e-40, -1.0, 9999.0, 10000.0, 12000.0]])
    if material == "iron_carbon":
        # embedding cludge threshold 28.5 and the rho cap 257
        g = np.concatenate([g, np.linspace(60.0, 400.0, 171), [28.5, np.nextafter(28.5, 99.0), 257.0,
                                                               np.nextafter(257.0, 999.0)]])
    return g


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
                  press=0.0, boxmass=50.0, flags=0, driver=DRIVER, pot_file=REF_POT,
                  material="fetransmetal", iverlet=1, nose=0.0, temprq=300.0):
    import refrun
    if iverlet == 0:
        flags |= 8   # PREDIC, FORCE, CORREC
    rg, rhog = r_grid(material), rho_grid(material)
    n = sysdef.n
    _, buf = refrun.run(sysdef, ivol=ivol, ntcm=ntcm, nsteps=nsteps, uselookup=uselookup,
                        nnbrs=nnbrs, rpad=rpad, press=press, boxmass=boxmass, flags=flags,
                        rgrid=rg, rhogrid=rhog, driver=driver, pot_file=pot_file, nose=nose, temprq=temprq)
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
    if flags & 32:
        d["pe_q"], d["vol_q"] = R.f8(2)
    if iverlet == 0:
        x3 = R.f8(3 * n).reshape(3, n)
        for k, nm in enumerate(["x3", "y3", "z3"]):
            d[nm + "_f"] = x3[k]
        d["b3_f"] = R.f8(9).reshape(3, 3, order="F")
    if flags & 4:
        R.f8(13 * max(nsteps, 1) + 4 * n)   # extra record (per-step trace, en_atom, x1)
    if flags & 128:
        d["embed_pe"] = float(R.f8(1)[0])
        d["embed_afrho"], d["embed_dafrho"] = R.f8(n), R.f8(n)
    if flags & 256:
        d["stress0"] = R.f8(3 * n).reshape(n, 3)
        d["stress_f"] = R.f8(3 * n).reshape(3, n).T.copy()
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
    # the force loop's per-atom diagonal stress (flag 256, atomstress) on the
    # sheared NPT box, the half list rebuilt every second force (ntcm=2)
    "atomstress_tric_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=97,
                                  strain=np.array([[0.01, 0.004, -0.003], [0.004, -0.006, 0.002],
                                                   [-0.003, 0.002, 0.003]])),
                             dict(ivol=0, ntcm=2, nsteps=5, uselookup=0, press=0.0, boxmass=50.0, flags=256)),
    # lookup tables (set_up before potential_set_lookup)
    "lookup_2000": (dict(nc=10, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=77),
                    dict(ivol=1, ntcm=10, nsteps=3, uselookup=1)),
    # hot dense liquid-like state: close approaches reach the Biersack/join region
    "hot_1024": (dict(nc=8, a=2.98118, sigma=0.30, temp=2500.0, deltat_fs=0.5, seed=5),
                 dict(ivol=1, ntcm=5, nsteps=10, uselookup=0)),
    # iron_carbon material (ref_driver_fec): PR NPT at 0.1 GPa on a sheared box, exact mode
    "fec_npt_1024": (dict(nc=8, a=2.87, sigma=0.05, temp=600.0, deltat_fs=1.0, seed=31,
                          strain=np.array([[0.008, 0.003, -0.002], [0.003, -0.004, 0.001],
                                           [-0.002, 0.001, 0.002]])),
                     dict(ivol=0, ntcm=10, nsteps=12, uselookup=0, press=0.1, boxmass=50.0,
                          material="iron_carbon")),
    # iron_carbon lookup tables (5000 points on [0, 5.4] A), NVE
    "fec_lookup_2000": (dict(nc=10, a=2.87, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=41),
                        dict(ivol=1, ntcm=10, nsteps=3, uselookup=1, material="iron_carbon")),
    # Gear predictor-corrector (iverlet=0): PR NPT on a sheared box, and the
    # z-only box (ivol=4)
    "gear_npt_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=61,
                           strain=np.array([[0.01, 0.004, -0.003], [0.004, -0.006, 0.002],
                                            [-0.003, 0.002, 0.003]])),
                      dict(ivol=0, ntcm=10, nsteps=12, uselookup=0, press=0.5, boxmass=50.0, iverlet=0)),
    "gear_z_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=62),
                    dict(ivol=4, ntcm=10, nsteps=12, uselookup=0, press=0.5, boxmass=50.0, iverlet=0)),
    # Nose thermostat as moldin drives it (flag 16): scalev about once per fs
    "nose_nvt_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=71),
                      dict(ivol=1, ntcm=10, nsteps=15, uselookup=0, flags=16, nose=50.0, temprq=600.0)),
    # quench (flag 32, the reference's VA14CD conjugate gradients): a 5^3 bcc
    # cell without its first atom, displaced by 0.1 A, box free at 0 GPa
    # embed on a caller-changed RHO after 3 steps (driver flag 128)
    "embed_rho_250": (dict(nc=5, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=82),
                      dict(ivol=1, ntcm=10, nsteps=3, uselookup=0, flags=128)),
    "quench_vac_249": (dict(nc=5, a=2.8665, sigma=0.1, temp=0.0, deltat_fs=1.0, seed=81, vacancy=0),
                       dict(ivol=0, ntcm=10, nsteps=0, uselookup=0, flags=32, press=0.0, boxmass=50.0)),
    # iron_carbon hot dense state: pairs inside the exp-cubic join (r < 2.05 A)
    "fec_hot_1024": (dict(nc=8, a=2.98118, sigma=0.30, temp=2500.0, deltat_fs=0.5, seed=43),
                     dict(ivol=1, ntcm=5, nsteps=10, uselookup=0, material="iron_carbon")),
    # open boundaries (pr_get_realsep_from_dx without an image along the open
    # axes): each open axis is centred on fractional 0, so the atoms straddle
    # the fold at 0, and widened 1.6x with the Cartesian positions kept.
    # ivol=2: free surfaces normal to z, the lowest eighth of the slab immobile (mass 0)
    "slab_ivol2_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=91,
                             open_axes=(2,), immobile=0.125),
                        dict(ivol=2, ntcm=10, nsteps=12, uselookup=0)),
    # the slab with a new run's velocities as moldin makes them (flag 2: the
    # reference's setvel, whose box part draws b1 for ivol=2, and scalev) and
    # the Nose thermostat during the run (flag 16)
    "setvel_slab_ivol2_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=0.0, deltat_fs=1.0, seed=94,
                                    open_axes=(2,), immobile=0.125),
                               dict(ivol=2, ntcm=10, nsteps=12, uselookup=0, flags=18, nose=50.0,
                                    temprq=300.0, boxmass=50.0)),
    # ivol=5: a pillar, periodic along z only
    "pillar_ivol5_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=92, open_axes=(0, 1)),
                          dict(ivol=5, ntcm=10, nsteps=12, uselookup=0)),
    # ivol=3: a free cluster, no periodic axis
    "cluster_ivol3_1024": (dict(nc=8, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=93,
                                open_axes=(0, 1, 2)),
                           dict(ivol=3, ntcm=10, nsteps=12, uselookup=0)),
}


def open_box(sd, axes, immobile=0.0, widen=1.6):
    """Open boundaries along axes: positions there centred on fractional 0 and
    the box widened by widen (Cartesian positions and velocities kept); with
    immobile > 0 the atoms of the lowest fraction along z get mass 0 and no
    velocity (the reference moves atoms with mass < 0.1 only by x1)."""
    b0, s, x1, mass = sd.b0.copy(), sd.s.copy(), sd.x1.copy(), sd.mass.copy()
    for ax in axes:
        s[:, ax] = (s[:, ax] - 0.5) / widen
        x1[:, ax] = x1[:, ax] / widen
        b0[:, ax] = b0[:, ax] * widen
    if immobile > 0.0:
        low = np.argsort(s[:, 2], kind="stable")[:int(round(immobile * sd.n))]
        mass[low] = 0.0
        x1[low] = 0.0
    return lattice.System(b0=b0, s=s, x1=x1, mass=mass, z=sd.z, deltat_fs=sd.deltat_fs)

KEEP_FIRST = ["rho0", "afrho0", "dafrho0", "fx0", "fy0", "fz0", "x20", "y20", "z20"]
KEEP_FINAL = ["x0_f", "y0_f", "z0_f", "x1_f", "y1_f", "z1_f", "fx_f", "fy_f", "fz_f"]
KEEP_GEAR = ["x2_f", "y2_f", "z2_f", "x3_f", "y3_f", "z3_f", "b3_f"]
KEEP_TENS = ["tp0", "tk0", "tc0", "fb0", "tgid0", "b20", "b0_f", "b1_f", "b2_f", "tp_f", "fb_f",
             "tgid_f", "tk_f"]


def _printed(tok: str):
    """A number as printed and half a unit in its last printed digit."""
    t = tok.upper().replace("D", "E")
    mant, _, ex = t.partition("E")
    dec = len(mant.split(".")[1]) if "." in mant else 0
    return float(t), 0.5 * 10.0 ** (-dec + (int(ex) if ex else 0))


def bcc_iron_fixture():
    """The reference's own shipped test tests/reference/bcc_iron: its inputs
    (params.in, 25^3.in) and what output66.txt and system.out print after 10
    steps, each value with half a unit of its last printed digit.  system.out
    is sampled every 25th atom (1250 atoms) with the full-precision box."""
    import re
    txt = open(os.path.join(BCC_IRON, "output66.txt")).read()
    fx = {}

    def grab(label, count, after=None):
        pos = txt.index(after) if after else 0
        m = re.search(re.escape(label) + r"((?:\s*[-+]?[0-9.]+(?:[EeDd][-+]?[0-9]+)?){%d})" % count,
                      txt[pos:])
        toks = m.group(1).split()
        vals = [_printed(t) for t in toks]
        return np.array([v for v, _ in vals]), np.array([h for _, h in vals])

    block = txt[txt.index("RUNNING AVERAGES AT STEP NO."):]
    for line in block.splitlines():
        m = re.match(r"^\s*(PE|KE|TE|T|H|TH|VOL):\s*(\S+)\s+A\w*:\s*(\S+)\s+\+/-\s*(\S+)", line)
        if not m or ("o66_" + m.group(1).lower()) in fx:
            continue
        key = m.group(1).lower()
        for pre, tok in (("", m.group(2)), ("a", m.group(3)), ("d", m.group(4))):
            v, h = _printed(tok)
            fx["o66_" + pre + key], fx["o66_" + pre + key + "_tol"] = np.float64(v), np.float64(h)
    st, sth = [], []
    for i in (1, 2, 3):
        v, h = grab("STRESS I= %d" % i, 3)
        st.append(v); sth.append(h)
    fx["o66_stress"], fx["o66_stress_tol"] = np.array(st), np.array(sth)
    mean, inst = [], []
    for i in (1, 2, 3):
        m = re.search(r"I=\s+%d MEAN VALUE:\s+(\S+)\s+(\S+)\s+(\S+)\s+INSTANTANEOUS VALUE:\s+(\S+)\s+(\S+)\s+(\S+)" % i, txt)
        mean.append([_printed(t) for t in m.groups()[:3]])
        inst.append([_printed(t) for t in m.groups()[3:]])
    # MEAN/INSTANTANEOUS rows print b0(:,i), column i of the box
    fx["o66_b0_mean"] = np.array([[v for v, _ in r] for r in mean]).T
    fx["o66_b0_mean_tol"] = np.array([[h for _, h in r] for r in mean]).T
    fx["o66_b0"] = np.array([[v for v, _ in r] for r in inst]).T
    fx["o66_b0_tol"] = np.array([[h for _, h in r] for r in inst]).T
    # the printed lines themselves, for the writers: PE..VOL of the averages
    # block and the six box lines
    keep = [ln for ln in block.splitlines() if re.match(r"^\s*(PE|KE|TE|T|H|TH|VOL):", ln)][:7]
    keep += [ln for ln in txt.splitlines() if "MEAN VALUE:" in ln or "INSTANTANEOUS VALUE:" in ln][:6]
    fx["o66_lines"] = np.array(keep)
    lines = open(os.path.join(BCC_IRON, "system.out")).read().splitlines()
    n = int(lines[0].split()[0])
    fx["sys_b0"] = np.array([[float(t.replace("D", "E")) for t in lines[2 + i].split()] for i in range(3)]).T
    idx = np.arange(0, n, 25)
    rows = [lines[5 + i].split() for i in idx]
    fx["sys_index"] = idx.astype(np.int32)
    fx["sys_pos_text"] = np.array([lines[5 + i][:42] for i in idx])
    for k, nm in enumerate(("x0", "y0", "z0")):
        vals = [_printed(r[k]) for r in rows]
        fx["sys_" + nm] = np.array([v for v, _ in vals])
        fx["sys_" + nm + "_tol"] = np.array([h for _, h in vals])
    vals = [_printed(r[5]) for r in rows]
    fx["sys_en"] = np.array([v for v, _ in vals])
    fx["sys_en_tol"] = np.array([h for _, h in vals])
    # chkp_out.dat: Fortran sequential records (4-byte markers) written by
    # write_checkpoint_file (src/moldin.F90:1053-1070); read as raw little-endian
    # numbers only.  Records: simparam, (x0,y0,z0,b0), (x1,y1,z1,b1),
    # (x2,y2,z2,b2), (x3,y3,z3,b3), atomic_number, mass, index, ispec, thmsums.
    raw = open(os.path.join(BCC_IRON, "chkp_out.dat"), "rb").read()
    recs, o = [], 0
    while o < len(raw):
        ln = int(np.frombuffer(raw, "<i4", 1, o)[0])
        recs.append(raw[o + 4:o + 4 + ln])
        o += ln + 8
    chk_s, chk_b = [], []
    for r in recs[1:5]:
        v = np.frombuffer(r, "<f8")
        chk_s.append(v[:3 * n].reshape(3, n)[:, idx])
        chk_b.append(v[3 * n:].reshape(3, 3, order="F"))
    fx["chk_s"] = np.concatenate(chk_s)          # rows x0 y0 z0 x1 y1 z1 x2 y2 z2 x3 y3 z3
    fx["chk_b"] = np.array(chk_b)                # b0, b1, b2, b3
    th = recs[9]
    fx["chk_thm"] = np.frombuffer(th[:192], "<f8").copy()   # SPE..SVOLPE, SB0(3,3)
    fx["chk_numsum"] = np.int64(np.frombuffer(th[192:196], "<i4")[0])
    fx["chk_z"] = np.frombuffer(recs[5], "<i4")[idx].copy()
    fx["chk_mass"] = np.frombuffer(recs[6], "<f8")[idx].copy()
    # inputs (params.in, 25^3.in)
    fx.update(n=np.int64(n), nt=np.array([25, 25, 25], dtype=np.int32), cell=np.diag([2.870] * 3),
              basis=np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]), basis_z=np.array([26, 26], dtype=np.int32),
              basis_mass=np.array([55.847, 55.847]), dsp=np.float64(0.0), temprq=np.float64(600.0),
              press_gpa=np.float64(0.1), boxmass=np.float64(50.0), deltat_fs=np.float64(1.0),
              rpad=np.float64(1.0), nsteps=np.int64(10), ntcm=np.int64(2000), ivol=np.int64(0),
              uselookup=np.int64(1), nnbrs=np.int64(150))
    path = os.path.join(OUT, "bcc_iron_golden.npz")
    np.savez_compressed(path, **fx)
    print(f"bcc_iron: {len(idx)} sampled atoms -> {path} ({os.path.getsize(path) // 1024} KiB)")


def shear_inputs_fixture():
    """The system file of the reference's shipped test tests/reference/shear
    (3480 atoms given one by one, 348 of them immobile), kept as text for the
    main program's runs on its inputs."""
    text = open(os.path.join(REF_TESTS, "shear", "system.in")).read()
    path = os.path.join(OUT, "shear_inputs.npz")
    np.savez_compressed(path, system_in=np.array(text))
    print(f"shear_inputs -> {path} ({os.path.getsize(path) // 1024} KiB)")


def potin_fixture():
    """potential.dat of the reference's potin (driver flag 64: the Fe-Fe pair,
    one species) as a (lines, 9) array: r, V, phi, V + phi/F, phi', V' at the
    hcp first-neighbour distance and the bcc, fcc and hcp cohesive energies."""
    import refrun
    sd = lattice.make_system(nc=3, a=2.8665, sigma=0.0, temp=0.0, seed=0)
    _, _, files = refrun.run_keep(sd, ("potential.dat",), driver=DRIVER, pot_file=REF_POT, flags=1 | 64,
                                  nsteps=0)
    rows = [[float(t) for t in ln.split()] for ln in files["potential.dat"].decode().splitlines() if ln.strip()]
    path = os.path.join(OUT, "potin_fe.npz")
    np.savez_compressed(path, table=np.array(rows))
    print(f"potin_fe: {len(rows)} lines -> {path}")


def main(names=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if not os.path.exists(DRIVER):
        raise SystemExit("build the reference driver first: oracle/build_ref.sh")
    if names and "bcc_iron" in names:
        bcc_iron_fixture()
    if names and "shear_inputs" in names:
        shear_inputs_fixture()
    if names and "potin_fe" in names:
        potin_fixture()
    for name, (skw, rkw) in CASES.items():
        if names and name not in names:
            continue
        skw = dict(skw)
        vac = skw.pop("vacancy", None)
        axes = skw.pop("open_axes", ())
        immobile = skw.pop("immobile", 0.0)
        sd = lattice.make_system(**skw)
        if axes:
            sd = open_box(sd, axes, immobile)
        if vac is not None:       # remove one atom
            keep = np.arange(sd.n) != vac
            sd = lattice.System(b0=sd.b0, s=sd.s[keep], x1=sd.x1[keep], mass=sd.mass[keep], z=sd.z[keep],
                                deltat_fs=sd.deltat_fs)
        drv = DRIVER_FEC if rkw.get("material") == "iron_carbon" else DRIVER
        d = run_reference(sd, driver=drv, **rkw)
        fx = {"b0": sd.b0, "s": sd.s, "x1": sd.x1, "mass": sd.mass, "z": sd.z,
              "deltat_fs": np.float64(sd.deltat_fs)}
        for k, v in rkw.items():
            fx["run_" + k] = np.asarray(v)
        for k in KEEP_FIRST + KEEP_FINAL + KEEP_TENS + (KEEP_GEAR if rkw.get("iverlet") == 0 else []):
            fx[k] = d[k]
        if "pe_q" in d:
            fx["pe_q"], fx["vol_q"] = np.float64(d["pe_q"]), np.float64(d["vol_q"])
        if "stress0" in d:
            fx["stress0"], fx["stress_f"] = d["stress0"], d["stress_f"]
        if "embed_pe" in d:
            fx["embed_pe"] = np.float64(d["embed_pe"])
            fx["embed_afrho"], fx["embed_dafrho"] = d["embed_afrho"], d["embed_dafrho"]
            fx["rho_f"], fx["dafrho_f"] = d["rho_f"], d["dafrho_f"]
        for k in ("nlc", "rnear", "deltat", "bdel2", "press", "vol0", "pe0", "trace", "numn0"):
            fx[k] = np.asarray(d[k])
        fx["numn0"] = d["numn0"].astype(np.int16)
        fx["npairs0"] = np.int64(len(d["pairs0"]))
        fx["pairs0_sha256"] = np.array(po.pairset_digest(d["pairs0"]))
        if name in ("c0_nve_2000", "lookup_2000", "fec_npt_1024", "fec_lookup_2000"):
            for k in ("rgrid", "rhogrid", "vee", "dvee", "phi", "dphi", "emb", "demb"):
                fx[k] = d[k]
        path = os.path.join(OUT, name + ".npz")
        np.savez_compressed(path, **fx)
        print(f"{name}: n={sd.n} pairs={len(d['pairs0'])} -> {path} "
              f"({os.path.getsize(path) // 1024} KiB)")


if __name__ == "__main__":
    # no arguments: every case; otherwise the named cases ("bcc_iron" for the shipped test)
    main(sys.argv[1:] or None)
