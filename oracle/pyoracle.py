"""ctypes binding of the oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path (moldy_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

# Fe-Fe coefficients of the published potential (pot_026_026.in format):
from moldy_amd.potential import FE_POTENTIAL  # noqa: E402

_lib = None


class OrcPot(C.Structure):
    _fields_ = [
        ("material", C.c_int),
        ("vcut", C.c_double * 13), ("vpar", C.c_double * 13), ("join", C.c_double * 4),
        ("pcut", C.c_double * 3), ("ppar", C.c_double * 3),
        ("acludge", C.c_double), ("rcludge", C.c_double), ("rx", C.c_double),
        ("nv", C.c_int), ("np", C.c_int),
        ("av", C.c_double * 8), ("rv", C.c_double * 8),
        ("ap", C.c_double * 4), ("rp", C.c_double * 4),
        ("bier_cut1", C.c_double), ("bier_cut2", C.c_double),
        ("sc", C.c_double * 4),
        ("phi_cut_r", C.c_double), ("phi_cut_v", C.c_double),
        ("use_lookup", C.c_int), ("nlookup", C.c_int),
        ("tab", C.c_void_p * 6),
        ("r_min", C.c_double), ("r_max", C.c_double), ("vp_range", C.c_double),
        ("vp_reciprange", C.c_double), ("vp_recipnlkp", C.c_double), ("vp_recipstep", C.c_double),
        ("rho_min", C.c_double), ("rho_max", C.c_double), ("e_range", C.c_double),
        ("e_reciprange", C.c_double), ("e_recipnlkp", C.c_double), ("e_recipstep", C.c_double),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_pot_init.argtypes = [C.POINTER(OrcPot), C.c_int, dp, dp, C.c_int, dp, dp]
        L.orc_pot_init_iron_carbon.argtypes = [C.POINTER(OrcPot), dp]
        L.orc_pot_range.argtypes = [C.POINTER(OrcPot)]
        L.orc_pot_range.restype = C.c_double
        L.orc_pot_build_lookup.argtypes = [C.POINTER(OrcPot), C.c_int]
        L.orc_pot_free.argtypes = [C.POINTER(OrcPot)]
        for nm in ("orc_vee", "orc_dvee", "orc_phi", "orc_dphi", "orc_emb", "orc_demb",
                   "orc_vee_src", "orc_dvee_src", "orc_phi_src", "orc_dphi_src"):
            f = getattr(L, nm)
            f.argtypes = [C.POINTER(OrcPot), C.c_double]
            f.restype = C.c_double
        L.orc_create.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp, dp, ip, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.POINTER(OrcPot)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        for nm in ("orc_linkup", "orc_relink", "orc_update_neighbourlist", "orc_force",
                   "orc_md_init", "orc_md_step"):
            getattr(L, nm).argtypes = [C.c_void_p]
            getattr(L, nm).restype = C.c_int
        L.orc_set_integrator.argtypes = [C.c_void_p, C.c_int]
        L.orc_scalev.argtypes = [C.c_void_p, C.c_double, C.c_double]
        for nm in ("orc_rhoset", "orc_embed", "orc_force_pairs", "orc_kinten", "orc_predic", "orc_correc"):
            getattr(L, nm).argtypes = [C.c_void_p]
            getattr(L, nm).restype = None
        L.orc_get_array.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_set_array.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_get_tensor.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_set_tensor.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_get_scalar.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_scalar.restype = C.c_double
        L.orc_get_int.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_ntc.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_pairs.argtypes = [C.c_void_p, ip, ip, C.c_long]
        L.orc_get_pairs.restype = C.c_long
        L.orc_get_numn.argtypes = [C.c_void_p, ip]
        L.orc_get_ic.argtypes = [C.c_void_p, ip]
        L.orc_get_chains.argtypes = [C.c_void_p, ip, ip]
        L.orc_bruteforce_pairs.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, C.c_double,
                                           ip, ip, C.c_long]
        L.orc_bruteforce_pairs.restype = C.c_long
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_eval_cell.argtypes = [C.POINTER(OrcPot), C.c_int, dp, dp, C.c_double, dp, dp, dp, dp]
        L.orc_energy_calc.argtypes = [C.c_void_p, dp]
        L.orc_rstart.argtypes = [C.POINTER(OrcRng), C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_rand1.argtypes = [C.POINTER(OrcRng)]
        L.orc_rand1.restype = C.c_double
        L.orc_replicate.argtypes = [C.POINTER(OrcRng), C.c_int, dp, dp, dp, ip, dp, C.c_double,
                                    dp, dp, dp]
        L.orc_setvel.argtypes = [C.POINTER(OrcRng), C.c_int, ip, dp, dp, C.c_double, C.c_double,
                                 dp, dp, dp]
        L.orc_setvel.restype = C.c_double
        L.orc_setvel_box.argtypes = [C.POINTER(OrcRng), C.c_int, C.c_int, C.c_double, C.c_double, dp]
        L.orc_setvel_box.restype = None
        L.orc_rqke.argtypes = [C.c_double, C.c_int]
        L.orc_rqke.restype = C.c_double
        L.orc_bk.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class OrcRng(C.Structure):
    _fields_ = [("u", C.c_double * 97), ("c", C.c_double), ("cd", C.c_double), ("cm", C.c_double),
                ("iu", C.c_int), ("ju", C.c_int)]


class Potential:
    """Fe-Fe potential of the oracle.  material "fetransmetal" (set_up_m
    constants computed at init, knot coefficients from coeffs) or
    "iron_carbon" (the module's coefficients, or the 38 iron.in values)."""

    def __init__(self, coeffs=None, lookup: bool = False, nlookup: int = 5000,
                 material: str = "fetransmetal", iron_in=None):
        self.s = OrcPot()
        self.material = material
        if material == "iron_carbon":
            self._keep = [np.ascontiguousarray(iron_in, dtype=np.float64)] if iron_in is not None else []
            lib().orc_pot_init_iron_carbon(C.byref(self.s), _dp(self._keep[0]) if self._keep else None)
        elif material == "fetransmetal":
            c = coeffs if coeffs is not None else FE_POTENTIAL
            self._keep = [np.ascontiguousarray(c[k], dtype=np.float64) for k in ("av", "rv", "ap", "rp")]
            av, rv, ap, rp = self._keep
            lib().orc_pot_init(C.byref(self.s), len(av), _dp(av), _dp(rv), len(ap), _dp(ap), _dp(rp))
        else:
            raise ValueError(material)
        if lookup:
            lib().orc_pot_build_lookup(C.byref(self.s), nlookup)

    def __del__(self):
        try:
            if self.s.use_lookup:
                lib().orc_pot_free(C.byref(self.s))
        except Exception:
            pass

    def f(self, name: str, x):
        fn = getattr(lib(), "orc_" + name)
        x = np.atleast_1d(np.asarray(x, dtype=np.float64))
        return np.array([fn(C.byref(self.s), float(v)) for v in x])

    @property
    def spline_coeff(self):
        return np.array(self.s.sc[:])

    @property
    def phi_cut(self):
        return (self.s.phi_cut_r, self.s.phi_cut_v)


SCALARS = {"pe": 0, "ke": 1, "vol": 2, "s1": 3, "s2": 4, "s3": 5, "te": 6, "h": 7,
           "th": 8, "boxke": 9, "deltat": 10, "bdel2": 11, "press": 12, "rnear": 13}
INTS = {"n": 0, "nlcx": 1, "nlcy": 2, "nlcz": 3, "ntc": 4, "ntcm": 5, "nnbrs": 6, "npairs": 7}


class OracleSystem:
    """Reference-semantics MD state on the CPU (serial unless threads>1)."""

    def __init__(self, sysdef, pot: Potential, ivol: int = 1, rpad: float = 0.3,
                 rcut: float = 0.0, nnbrs: int = 150, ntcm: int = 10,
                 press_gpa: float = 0.0, boxmass: float = 50.0, iverlet: int = 1):
        self.pot = pot
        self.ivol = int(ivol)
        n = sysdef.n
        b0 = np.asfortranarray(sysdef.b0).ravel(order="F").copy()
        s = np.ascontiguousarray(sysdef.s.T)
        x1 = np.ascontiguousarray(sysdef.x1.T)
        self._z = np.ascontiguousarray(sysdef.z, dtype=np.int32)
        mass = np.ascontiguousarray(sysdef.mass, dtype=np.float64)
        self.n = n
        self.h = lib().orc_create(n, _dp(b0), _dp(s[0]), _dp(s[1]), _dp(s[2]), _dp(x1[0]),
                                  _dp(x1[1]), _dp(x1[2]), _dp(mass), _ip(self._z), ivol,
                                  sysdef.deltat_fs, rpad, rcut, nnbrs, ntcm, press_gpa,
                                  boxmass, C.byref(pot.s))
        if not self.h:
            raise MemoryError("orc_create failed")
        lib().orc_set_integrator(self.h, int(iverlet))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def call(self, name: str) -> int:
        r = getattr(lib(), "orc_" + name)(self.h)
        return 0 if r is None else int(r)

    def array(self, name: str) -> np.ndarray:
        out = np.empty(self.n)
        if lib().orc_get_array(self.h, name.encode(), _dp(out)) != 0:
            raise KeyError(name)
        return out

    def set_array(self, name: str, v) -> None:
        v = np.ascontiguousarray(v, dtype=np.float64)
        if lib().orc_set_array(self.h, name.encode(), _dp(v)) != 0:
            raise KeyError(name)

    def tensor(self, name: str) -> np.ndarray:
        out = np.empty(9)
        if lib().orc_get_tensor(self.h, name.encode(), _dp(out)) != 0:
            raise KeyError(name)
        return out.reshape(3, 3, order="F")

    def set_tensor(self, name: str, t) -> None:
        t = np.asfortranarray(t, dtype=np.float64).ravel(order="F").copy()
        lib().orc_set_tensor(self.h, name.encode(), _dp(t))

    def scalar(self, name: str) -> float:
        return lib().orc_get_scalar(self.h, SCALARS[name])

    def int(self, name: str) -> int:
        return lib().orc_get_int(self.h, INTS[name])

    def pairs(self) -> np.ndarray:
        cnt = lib().orc_get_pairs(self.h, None, None, 0)
        pi = np.empty(max(cnt, 1), dtype=np.int32)
        pj = np.empty(max(cnt, 1), dtype=np.int32)
        lib().orc_get_pairs(self.h, _ip(pi), _ip(pj), cnt)
        return np.stack([pi[:cnt], pj[:cnt]], axis=1)

    def numn(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        lib().orc_get_numn(self.h, _ip(out))
        return out

    def ic(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        lib().orc_get_ic(self.h, _ip(out))
        return out

    def chains(self):
        """(L0, LINK) of the link cells, 1-based, as linkup/relink leave them."""
        m = self.int("nlcx") * self.int("nlcy") * self.int("nlcz")
        l0 = np.empty(m, dtype=np.int32)
        link = np.empty(self.n, dtype=np.int32)
        lib().orc_get_chains(self.h, _ip(l0), _ip(link))
        return l0, link

    def scalev(self, nose: float, rqke: float) -> None:
        lib().orc_scalev(self.h, float(nose), float(rqke))

    def box_quench(self) -> None:
        """moldin's iquen = 2 step (src/moldin.F90:845-853): each b1(i,j) is
        kept only when b1(i,j) * b2(i,j) > 0, else set to zero."""
        b1, b2 = self.tensor("b1"), self.tensor("b2")
        for i in range(3):
            for j in range(3):
                if not b1[i, j] * b2[i, j] > 0.0:
                    b1[i, j] = 0.0
        self.set_tensor("b1", b1)

    def energy_calc(self) -> np.ndarray:
        """en_atom of energy_calc (needs afrho of a previous force())."""
        out = np.empty(self.n)
        lib().orc_energy_calc(self.h, _dp(out))
        return out

    def atomic_stress(self) -> np.ndarray:
        """(n, 3) per-atom diagonal stress of the force loop's atomstress
        option (src/dynamics_m.F90:279-283): stress_x(i) = sum of rx_ij fp_ij
        over atom i's own half-list partners j, rx = B0 (s_i - s_j) after
        pr_get_realsep_from_dx's minimum image, fp = -V'/r - phi'(F'_i + F'_j)/r.
        Needs the list and F' of a previous force()."""
        p = self.pairs()
        i, j = p[:, 0], p[:, 1]
        s = np.stack([self.array("x0"), self.array("y0"), self.array("z0")], axis=1)
        d = s[i] - s[j]
        per = {0: (0, 1, 2), 1: (0, 1, 2), 4: (0, 1, 2), 2: (0, 1), 5: (2,)}.get(self.ivol, ())
        for k in per:
            x = d[:, k]
            d[:, k] = x - np.where(x >= 0, np.floor(x + 0.5), np.ceil(x - 0.5))   # nint
        rvec = d @ self.tensor("b0").T
        r = np.sqrt(np.sum(rvec * rvec, axis=1))
        df = self.array("dafrho")
        fp = -self.pot.f("dvee", r) / r - self.pot.f("dphi", r) * (df[i] + df[j]) / r
        out = np.zeros((self.n, 3))
        np.add.at(out, i, rvec * fp[:, None])
        return out

    def forces_cart(self) -> np.ndarray:
        f = np.stack([self.array("fx"), self.array("fy"), self.array("fz")], axis=1)
        return f @ self.tensor("b0").T


class Ranmar:
    """The reference's RANMAR stream (rstart / rand1)."""

    def __init__(self, i=97, j=33, k=21, l=82):
        self.g = OrcRng()
        lib().orc_rstart(C.byref(self.g), i, j, k, l)

    def rand1(self, count: int = 1) -> np.ndarray:
        return np.array([lib().orc_rand1(C.byref(self.g)) for _ in range(count)])


def bk() -> float:
    return lib().orc_bk()


def moldin_new_run(basis_frac, basis_z, basis_mass, nt, cell_b0, dsp: float, temprq: float,
                   deltat_fs: float, ivol=None, boxmass: float = 0.0):
    """Initial state of a new run as moldin makes it: rstart(97,33,21,82),
    read_system's replication (with its 3 draws per atom) and setvel.
    Returns (b0 of the whole box, s (n,3), x1 (n,3), z, mass); with ivol given,
    also setvel's box velocities b1 (3,3) for that ivol and box mass per atom."""
    g = Ranmar()
    nb = len(basis_z)
    nt = np.ascontiguousarray(nt, dtype=np.int32)
    b0 = np.asarray(cell_b0, dtype=np.float64) * nt[None, :]   # b0(:,i) *= nt(i)
    bf = np.asfortranarray(b0).ravel(order="F").copy()
    xb = np.ascontiguousarray(np.asarray(basis_frac, dtype=np.float64).T)
    n = nb * int(np.prod(nt))
    x0, y0, z0 = np.empty(n), np.empty(n), np.empty(n)
    lib().orc_replicate(C.byref(g.g), nb, _dp(xb[0]), _dp(xb[1]), _dp(xb[2]), _ip(nt), _dp(bf), dsp,
                        _dp(x0), _dp(y0), _dp(z0))
    z = np.ascontiguousarray(np.tile(np.asarray(basis_z, dtype=np.int32), n // nb))
    mass = np.ascontiguousarray(np.tile(np.asarray(basis_mass, dtype=np.float64), n // nb))
    x1, y1, z1 = np.zeros(n), np.zeros(n), np.zeros(n)
    # deltat in model units: fs * fs_to_time as read_params converts it
    from moldy_amd import lattice
    dt = deltat_fs * lattice.FS_TO_TIME
    lib().orc_setvel(C.byref(g.g), n, _ip(z), _dp(mass), _dp(bf), dt, lib().orc_rqke(temprq, n),
                     _dp(x1), _dp(y1), _dp(z1))
    out = (b0, np.stack([x0, y0, z0], axis=1), np.stack([x1, y1, z1], axis=1), z, mass)
    if ivol is None:
        return out
    b1 = np.zeros(9)
    lib().orc_setvel_box(C.byref(g.g), int(ivol), n, float(temprq), dt * dt / (2.0 * n * boxmass), _dp(b1))
    return out + (b1.reshape(3, 3, order="F"),)


def new_run_velocities(sysdef, temprq: float, ivol: int, boxmass: float):
    """Velocities of a new run for a system given atom by atom, as moldin
    makes them (and the reference driver with its flag 2): rstart(97,33,21,82),
    the 3 draws per atom of read_system, setvel's atom part with
    RQKE = 1.5 bk temprq (nm - 1), then its box part (ivol 1, 2, 3, 5).
    Returns (x1 (n,3), b1 (3,3)); atoms that setvel skips keep sysdef.x1."""
    g = Ranmar()
    n = sysdef.n
    for _ in range(3 * n):
        lib().orc_rand1(C.byref(g.g))
    z = np.ascontiguousarray(sysdef.z, dtype=np.int32)
    mass = np.ascontiguousarray(sysdef.mass, dtype=np.float64)
    bf = np.asfortranarray(sysdef.b0).ravel(order="F").copy()
    x1 = np.ascontiguousarray(np.asarray(sysdef.x1, dtype=np.float64).T).copy()
    from moldy_amd import lattice
    dt = sysdef.deltat_fs * lattice.FS_TO_TIME
    lib().orc_setvel(C.byref(g.g), n, _ip(z), _dp(mass), _dp(bf), dt, lib().orc_rqke(temprq, n),
                     _dp(x1[0]), _dp(x1[1]), _dp(x1[2]))
    b1 = np.zeros(9)
    bdel2 = dt * dt / (2.0 * n * boxmass)
    lib().orc_setvel_box(C.byref(g.g), int(ivol), n, float(temprq), bdel2, _dp(b1))
    return x1.T.copy(), b1.reshape(3, 3, order="F")


def bruteforce_pairs(b0, s, rcut: float, ivol: int = 1) -> np.ndarray:
    b = np.asfortranarray(b0).ravel(order="F").copy()
    st = np.ascontiguousarray(np.asarray(s).T)
    n = st.shape[1]
    cnt = lib().orc_bruteforce_pairs(n, _dp(b), _dp(st[0]), _dp(st[1]), _dp(st[2]), ivol,
                                     rcut, None, None, 0)
    pi = np.empty(max(cnt, 1), dtype=np.int32)
    pj = np.empty(max(cnt, 1), dtype=np.int32)
    lib().orc_bruteforce_pairs(n, _dp(b), _dp(st[0]), _dp(st[1]), _dp(st[2]), ivol, rcut,
                               _ip(pi), _ip(pj), cnt)
    return np.stack([pi[:cnt], pj[:cnt]], axis=1)


def canonical_pairs(p: np.ndarray) -> np.ndarray:
    """Unordered pair set as sorted (min,max) rows."""
    if len(p) == 0:
        return np.zeros((0, 2), dtype=np.int64)
    q = np.sort(np.asarray(p, dtype=np.int64), axis=1)
    key = q[:, 0] * (1 << 32) + q[:, 1]
    order = np.argsort(key, kind="stable")
    return q[order]


def pairset_digest(p: np.ndarray) -> str:
    import hashlib
    q = canonical_pairs(p).astype("<i8")
    return hashlib.sha256(q.tobytes()).hexdigest()


def eval_cell(pot: Potential, sysdef, rcut: float = 3.72645 + 1e-9):
    """(pe, Cartesian forces (n,3), rho (n,), W (3,3)) of one periodic cell, brute force."""
    n = sysdef.n
    b = np.asfortranarray(sysdef.b0).ravel(order="F").copy()
    s = np.ascontiguousarray(sysdef.s, dtype=np.float64)
    pe = C.c_double()
    f = np.empty(3 * n)
    rho = np.empty(n)
    w = np.empty(9)
    lib().orc_eval_cell(C.byref(pot.s), n, _dp(b), _dp(s), rcut, C.byref(pe), _dp(f), _dp(rho), _dp(w))
    return pe.value, f.reshape(n, 3), rho, w.reshape(3, 3, order="F")
