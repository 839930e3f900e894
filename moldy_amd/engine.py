"""ctypes binding of libmoldy_hip.so (the C ABI declared in include/moldy_hip.h).

This is the Python mirror of the reference's Fortran module procedures for the
hot path (linkcell_m, neighbourlist_m, potential_m, dynamics_m,
parrinellorahman_m, quench_m::velocityverlet).  Host arrays are numpy arrays in
the reference's atom order.  There is no CPU fallback: when the HIP library is
missing or no GPU is visible every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import lattice
from .potential import FE_POTENTIAL

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmoldy_hip.so")

MDK_OK, MDK_ERR_HIP, MDK_ERR_NBR_OVERFLOW, MDK_ERR_BAD_ARG = 0, 1, 2, 3
MDK_ERR_SINGULAR_BOX, MDK_ERR_STATE, MDK_ERR_SKIN, MDK_ERR_NO_DEVICE = 4, 5, 6, 7
PREC_FP64, PREC_FP32 = 0, 1

FIELDS = ["x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2", "fx", "fy", "fz",
          "rho", "afrho", "dafrho", "mass", "del2", "x3", "y3", "z3"]
TENSORS = ["b0", "b1", "b2", "tk", "tc", "tg", "tp", "tginv", "tgdot", "tgid", "fb", "b3"]
SCALARS = ["pe", "ke", "vol", "s1", "s2", "s3", "te", "h", "th", "boxke", "maxdisp"]

# Every symbol include/moldy_hip.h declares (checked by the CPU test suite).
EXPORTS = [
    "mdk_create", "mdk_destroy", "mdk_last_error", "mdk_set_potential", "mdk_set_lookup",
    "mdk_potential_constants", "mdk_eval_potential", "mdk_upload", "mdk_download",
    "mdk_upload_species", "mdk_set_tensor", "mdk_get_tensor", "mdk_get_scalars", "mdk_set_ntc",
    "mdk_get_ntc", "mdk_linkup", "mdk_relink", "mdk_get_linkcells", "mdk_update_neighbourlist",
    "mdk_get_neighbourlist", "mdk_get_pairs", "mdk_force", "mdk_rhoset", "mdk_embed",
    "mdk_kinten", "mdk_md_init", "mdk_velocityverlet", "mdk_update_therm_avs", "mdk_md_steps",
    "mdk_stats", "mdk_timing", "mdk_timing_read", "mdk_synchronize",
    "mdk_batch_create", "mdk_batch_destroy", "mdk_batch_last_error", "mdk_batch_set_potential",
    "mdk_batch_config", "mdk_batch_evaluate", "mdk_batch_synchronize", "mdk_batch_results",
    "mdk_batch_tune", "mdk_batch_time", "mdk_batch_info",
    "mdk_timing_every", "mdk_batch_set_targets", "mdk_batch_loss_grad", "mdk_batch_vjp", "mdk_batch_time_loss_grad",
    "mdk_batch_loss_grad_noalloc",
    "mdk_set_iron_carbon", "mdk_energy_calc", "mdk_atomic_stress", "mdk_set_integrator", "mdk_predic", "mdk_correc",
    "mdk_set_pressure", "mdk_scalev", "mdk_get_snhv", "mdk_set_snhv",
]

# get_supported_potential_range's rmax of each material module
# (src/fetransmetal.F90:589-593, src/iron_carbon.F90:93)
MATERIAL_RANGE = {"fetransmetal": 6.0, "iron_carbon": 5.4}


class MdkParams(C.Structure):
    _fields_ = [("nm", C.c_int), ("ivol", C.c_int), ("nnbrs", C.c_int), ("ntcm", C.c_int),
                ("nlcx", C.c_int), ("nlcy", C.c_int), ("nlcz", C.c_int),
                ("precision", C.c_int), ("rebuild", C.c_int), ("max_full", C.c_int),
                ("deltat", C.c_double), ("rnear", C.c_double), ("rskin", C.c_double),
                ("press", C.c_double), ("bdel2", C.c_double)]


class MdkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdk error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads the HIP library; raises when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libmoldy_hip.so not built at {path}: run __graft_entry__.build()")
    L = C.CDLL(path)
    vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
    sig = {
        "mdk_create": [C.POINTER(C.c_void_p), C.POINTER(MdkParams)],
        "mdk_destroy": [vp],
        "mdk_set_potential": [vp, C.c_int, dp, dp, C.c_int, dp, dp],
        "mdk_set_lookup": [vp, C.c_int],
        "mdk_set_iron_carbon": [vp, dp],
        "mdk_energy_calc": [vp, dp],
        "mdk_atomic_stress": [vp, dp, dp, dp],
        "mdk_set_integrator": [vp, C.c_int], "mdk_predic": [vp], "mdk_correc": [vp],
        "mdk_set_pressure": [vp, C.c_double],
        "mdk_scalev": [vp, C.c_double, C.c_double], "mdk_get_snhv": [vp, dp], "mdk_set_snhv": [vp, C.c_double],
        "mdk_potential_constants": [vp, dp],
        "mdk_eval_potential": [vp, C.c_int, dp, dp, C.c_int],
        "mdk_upload": [vp, C.c_int, dp],
        "mdk_download": [vp, C.c_int, dp],
        "mdk_upload_species": [vp, ip],
        "mdk_set_tensor": [vp, C.c_int, dp],
        "mdk_get_tensor": [vp, C.c_int, dp],
        "mdk_get_scalars": [vp, dp],
        "mdk_set_ntc": [vp, C.c_int],
        "mdk_get_ntc": [vp, ip],
        "mdk_linkup": [vp], "mdk_relink": [vp],
        "mdk_get_linkcells": [vp, ip, ip, ip],
        "mdk_update_neighbourlist": [vp],
        "mdk_get_neighbourlist": [vp, ip, ip],
        "mdk_get_pairs": [vp, lp, ip, ip, C.c_long],
        "mdk_force": [vp], "mdk_rhoset": [vp], "mdk_embed": [vp, dp], "mdk_kinten": [vp],
        "mdk_md_init": [vp], "mdk_velocityverlet": [vp],
        "mdk_update_therm_avs": [vp, dp, dp, dp],
        "mdk_md_steps": [vp, C.c_int, dp],
        "mdk_stats": [vp, lp, lp],
        "mdk_timing": [vp, C.c_int], "mdk_timing_every": [vp, C.c_int],
        "mdk_timing_read": [vp, dp, lp, C.c_int],
        "mdk_synchronize": [vp],
    }
    sig.update({
        "mdk_batch_create": [C.POINTER(C.c_void_p), C.c_int, ip, dp, dp, C.c_int],
        "mdk_batch_destroy": [vp],
        "mdk_batch_set_potential": [vp, C.c_int, dp, dp, C.c_int, dp, dp],
        "mdk_batch_config": [vp, C.c_int, C.c_int, C.c_int],
        "mdk_batch_evaluate": [vp], "mdk_batch_synchronize": [vp],
        "mdk_batch_results": [vp, dp, dp, dp, dp],
        "mdk_batch_tune": [vp, C.c_int, ip, ip, dp],
        "mdk_batch_time": [vp, C.c_int, dp],
        "mdk_batch_info": [vp, ip, ip, ip, lp],
        "mdk_batch_set_targets": [vp, dp, dp],
        "mdk_batch_loss_grad": [vp, dp, dp],
        "mdk_batch_vjp": [vp, dp, dp, dp],
        "mdk_batch_time_loss_grad": [vp, C.c_int, dp],
        "mdk_batch_loss_grad_noalloc": [vp, vp, vp],
    })
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    for nm in ("mdk_last_error", "mdk_batch_last_error"):
        getattr(L, nm).argtypes = []
        getattr(L, nm).restype = C.c_char_p
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc):
    if rc != MDK_OK:
        raise MdkError(rc, load_library().mdk_last_error().decode(errors="replace"))


FS_TO_TIME = lattice.FS_TO_TIME
GPA_TO_PRESS = 1.0 / ((1.602176487e-19 / (1e-10 * 1e-10 * 1e-10)) / 1.0e9)


def link_cells_for(b0: np.ndarray, rnear: float) -> tuple[int, int, int]:
    """moldin.F90:252-257: int(b0(i,i)/rnear), at least 3."""
    return tuple(max(3, int(b0[i, i] / rnear)) for i in range(3))


class Engine:
    """One simulated periodic cell on the GPU (reference semantics)."""

    def __init__(self, sysdef: lattice.System, ivol: int = 1, rpad: float = 0.3,
                 rcut: float = 0.0, nnbrs: int = 150, ntcm: int = 10, press_gpa: float = 0.0,
                 boxmass: float = 50.0, precision: int = PREC_FP64, rebuild: int = 0,
                 skin: float | None = None, max_full: int = 0, coeffs: dict | None = None,
                 lookup: int = 0, material: str = "fetransmetal", iron_in=None, iverlet: int = 1):
        """material "fetransmetal" (knot coefficients coeffs, default the Fe-Fe
        file) or "iron_carbon" (the module's coefficients, or the 38 values of
        an iron.in given as iron_in).  iverlet: 1 velocity Verlet, 0 Gear
        predictor-corrector for md_steps."""
        L = load_library()
        self.n = sysdef.n
        n = self.n
        if material not in MATERIAL_RANGE:
            raise ValueError(f"unknown material {material!r}")
        self.material = material
        rcut = max(rcut, MATERIAL_RANGE[material])  # moldin.F90:226 with the material's range
        self.rnear = rcut + rpad
        dt = sysdef.deltat_fs * FS_TO_TIME
        bmass = n * boxmass
        p = MdkParams()
        p.nm, p.ivol, p.nnbrs, p.ntcm = n, ivol, nnbrs, ntcm
        p.nlcx, p.nlcy, p.nlcz = link_cells_for(sysdef.b0, self.rnear)
        p.precision, p.rebuild, p.max_full = precision, rebuild, max_full
        p.deltat, p.rnear = dt, self.rnear
        p.rskin = skin if skin is not None else rpad
        p.press = press_gpa * GPA_TO_PRESS
        p.bdel2 = dt * dt / (2.0 * bmass)
        self.params = p
        h = C.c_void_p()
        _check(L.mdk_create(C.byref(h), C.byref(p)))
        self.h = h
        self.L = L
        if material == "iron_carbon":
            self._iron_in = None if iron_in is None else np.ascontiguousarray(iron_in, dtype=np.float64)
            if self._iron_in is not None and self._iron_in.shape != (38,):
                raise ValueError("iron_in holds 38 values")
            _check(L.mdk_set_iron_carbon(h, _dp(self._iron_in) if self._iron_in is not None else None))
        else:
            c = coeffs if coeffs is not None else FE_POTENTIAL
            self.set_potential(c)
        if lookup:
            _check(L.mdk_set_lookup(h, int(lookup)))
        s = np.ascontiguousarray(sysdef.s)
        x1 = np.ascontiguousarray(sysdef.x1)
        for k, nm in enumerate(("x0", "y0", "z0")):
            self.upload(nm, s[:, k])
        for k, nm in enumerate(("x1", "y1", "z1")):
            self.upload(nm, x1[:, k])
        zero = np.zeros(n)
        for nm in ("x2", "y2", "z2"):
            self.upload(nm, zero)
        mass = np.ascontiguousarray(sysdef.mass, dtype=np.float64)
        self.upload("mass", mass)
        # moldin.F90:294-300 (0.01 is a REAL(4) literal)
        del2 = np.where(mass > float(np.float32(0.01)), dt * dt / (2.0 * mass), 0.0)
        self.upload("del2", del2)
        _check(L.mdk_upload_species(h, _ip(np.ascontiguousarray(sysdef.z, dtype=np.int32))))
        self.set_tensor("b0", sysdef.b0)
        _check(L.mdk_set_integrator(h, int(iverlet)))

    def close(self):
        if getattr(self, "h", None):
            self.L.mdk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- potential
    def set_potential(self, c: dict):
        av, rv, ap, rp = (np.ascontiguousarray(c[k], dtype=np.float64) for k in ("av", "rv", "ap", "rp"))
        _check(self.L.mdk_set_potential(self.h, len(av), _dp(av), _dp(rv), len(ap), _dp(ap), _dp(rp)))

    def potential_constants(self) -> np.ndarray:
        out = np.empty(8)
        _check(self.L.mdk_potential_constants(self.h, _dp(out)))
        return out

    def eval_potential(self, which: str, x) -> np.ndarray:
        # vee32, dvee32, dphi32: the FP32 force pass's pair terms (interval table)
        k = ["vee", "dvee", "phi", "dphi", "emb", "demb", "vee32", "dvee32", "dphi32"].index(which)
        x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
        y = np.empty_like(x)
        _check(self.L.mdk_eval_potential(self.h, k, _dp(x), _dp(y), len(x)))
        return y

    # --- state
    def upload(self, name: str, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == (self.n,)
        _check(self.L.mdk_upload(self.h, FIELDS.index(name), _dp(v)))

    def array(self, name: str) -> np.ndarray:
        out = np.empty(self.n)
        _check(self.L.mdk_download(self.h, FIELDS.index(name), _dp(out)))
        return out

    def set_tensor(self, name: str, t):
        t = np.asfortranarray(np.asarray(t, dtype=np.float64)).ravel(order="F").copy()
        _check(self.L.mdk_set_tensor(self.h, TENSORS.index(name), _dp(t)))

    def tensor(self, name: str) -> np.ndarray:
        out = np.empty(9)
        _check(self.L.mdk_get_tensor(self.h, TENSORS.index(name), _dp(out)))
        return out.reshape(3, 3, order="F")

    def scalars(self) -> dict:
        out = np.empty(len(SCALARS))
        _check(self.L.mdk_get_scalars(self.h, _dp(out)))
        return dict(zip(SCALARS, out.tolist()))

    def scalar(self, name: str) -> float:
        return self.scalars()[name]

    @property
    def ntc(self) -> int:
        v = C.c_int()
        _check(self.L.mdk_get_ntc(self.h, C.byref(v)))
        return v.value

    @ntc.setter
    def ntc(self, v: int):
        _check(self.L.mdk_set_ntc(self.h, int(v)))

    # --- reference procedures
    def call(self, name: str):
        _check(getattr(self.L, "mdk_" + name)(self.h))

    def linkup(self):
        self.call("linkup")

    def relink(self):
        self.call("relink")

    def update_neighbourlist(self):
        self.call("update_neighbourlist")

    def force(self):
        self.call("force")

    def md_init(self):
        self.call("md_init")

    def set_pressure(self, press_model: float):
        """External pressure in model units (eV/A^3)."""
        _check(self.L.mdk_set_pressure(self.h, float(press_model)))
        self.params.press = float(press_model)

    def scalev(self, nose: float, rqke: float):
        """Nose-Hoover velocity scaling (scalev); rqke in eV."""
        _check(self.L.mdk_scalev(self.h, float(nose), float(rqke)))

    @property
    def snhv(self) -> float:
        v = C.c_double()
        _check(self.L.mdk_get_snhv(self.h, C.byref(v)))
        return v.value

    def predic(self):
        self.call("predic")

    def correc(self):
        self.call("correc")

    def energy_calc(self) -> np.ndarray:
        """Per-atom energies of energy_calc (reference atom order)."""
        out = np.empty(self.n)
        _check(self.L.mdk_energy_calc(self.h, _dp(out)))
        return out

    def atomic_stress(self) -> np.ndarray:
        """(n, 3) per-atom diagonal stress of force's atomstress option
        (reference atom order)."""
        out = np.empty((3, self.n))
        _check(self.L.mdk_atomic_stress(self.h, _dp(out[0]), _dp(out[1]), _dp(out[2])))
        return out.T.copy()

    def velocityverlet(self):
        self.call("velocityverlet")

    def update_therm_avs(self):
        s1, s2, s3 = C.c_double(), C.c_double(), C.c_double()
        _check(self.L.mdk_update_therm_avs(self.h, C.byref(s1), C.byref(s2), C.byref(s3)))
        return s1.value, s2.value, s3.value

    def md_steps(self, nsteps: int, trace: bool = True) -> np.ndarray | None:
        tr = np.empty((max(nsteps, 1), 6)) if trace else None
        _check(self.L.mdk_md_steps(self.h, int(nsteps), _dp(tr) if trace else None))
        return tr[:nsteps] if trace else None

    def linkcells(self):
        m = self.params.nlcx * self.params.nlcy * self.params.nlcz
        l0 = np.empty(m, dtype=np.int32)
        link = np.empty(self.n, dtype=np.int32)
        ic = np.empty(self.n, dtype=np.int32)
        _check(self.L.mdk_get_linkcells(self.h, _ip(l0), _ip(link), _ip(ic)))
        return l0, link, ic

    def neighbourlist(self):
        nn = self.params.nnbrs
        numn = np.empty(self.n, dtype=np.int32)
        nlist = np.empty(self.n * nn, dtype=np.int32)
        _check(self.L.mdk_get_neighbourlist(self.h, _ip(numn), _ip(nlist)))
        return numn, nlist.reshape(self.n, nn)

    def pairs(self) -> np.ndarray:
        cnt = C.c_long()
        _check(self.L.mdk_get_pairs(self.h, C.byref(cnt), None, None, 0))
        k = cnt.value
        pi = np.empty(max(k, 1), dtype=np.int32)
        pj = np.empty(max(k, 1), dtype=np.int32)
        _check(self.L.mdk_get_pairs(self.h, C.byref(cnt), _ip(pi), _ip(pj), k))
        return np.stack([pi[:k], pj[:k]], axis=1)

    def stats(self) -> tuple[int, int]:
        a, b = C.c_long(), C.c_long()
        _check(self.L.mdk_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    TIMING_CLASSES = ("drift", "density", "force", "kick", "rebuild")

    def timing(self, enable: bool, classes=None, every: int = 1):
        """Record HIP events around the kernels of the given classes (all by
        default; two events per recorded launch, one launch in `every` per
        class) or stop recording."""
        _check(self.L.mdk_timing_every(self.h, int(every)))
        mask = 0
        if enable:
            names = self.TIMING_CLASSES if classes is None else classes
            for nm in names:
                mask |= 1 << self.TIMING_CLASSES.index(nm)
        _check(self.L.mdk_timing(self.h, mask))

    def timing_read(self):
        ms = np.zeros(5)
        cnt = np.zeros(5, dtype=np.int64)
        _check(self.L.mdk_timing_read(self.h, _dp(ms), cnt.ctypes.data_as(C.POINTER(C.c_long)), 5))
        return dict(zip(["drift", "density", "force", "kick", "rebuild"], zip(ms.tolist(), cnt.tolist())))

    def synchronize(self):
        self.call("synchronize")

    def forces_cart(self) -> np.ndarray:
        f = np.stack([self.array("fx"), self.array("fy"), self.array("fz")], axis=1)
        return f @ self.tensor("b0").T


def _bcheck(rc):
    if rc != MDK_OK:
        raise MdkError(rc, load_library().mdk_batch_last_error().decode(errors="replace"))


class BatchEngine:
    """Many independent periodic cells evaluated in one persistent launch."""

    def __init__(self, cells: list, precision: int = PREC_FP32, coeffs: dict | None = None):
        L = load_library()
        self.L = L
        self.ncell = len(cells)
        self.natoms = np.array([c.n for c in cells], dtype=np.int32)
        self.boxes = np.concatenate([np.asarray(c.b0, dtype=np.float64).ravel(order="F") for c in cells])
        s = np.ascontiguousarray(np.concatenate([c.s for c in cells]), dtype=np.float64)
        self.natom = int(self.natoms.sum())
        h = C.c_void_p()
        _bcheck(L.mdk_batch_create(C.byref(h), self.ncell, _ip(self.natoms), _dp(self.boxes), _dp(s),
                                   int(precision)))
        self.h = h
        self.set_potential(coeffs if coeffs is not None else FE_POTENTIAL)

    def set_potential(self, c: dict):
        av, rv, ap, rp = (np.ascontiguousarray(c[k], dtype=np.float64) for k in ("av", "rv", "ap", "rp"))
        _bcheck(self.L.mdk_batch_set_potential(self.h, len(av), _dp(av), _dp(rv), len(ap), _dp(ap), _dp(rp)))
        self.ngrad = 2 * len(av) + 2 * len(ap)

    def config(self, block=256, wg_per_cu=4, graph=True):
        _bcheck(self.L.mdk_batch_config(self.h, block, wg_per_cu, 1 if graph else 0))

    def evaluate(self):
        _bcheck(self.L.mdk_batch_evaluate(self.h))

    def synchronize(self):
        _bcheck(self.L.mdk_batch_synchronize(self.h))

    def results(self):
        e = np.empty(self.ncell)
        w = np.empty(6 * self.ncell)
        f = np.empty(3 * self.natom)
        rho = np.empty(self.natom)
        _bcheck(self.L.mdk_batch_results(self.h, _dp(e), _dp(w), _dp(f), _dp(rho)))
        return e, w.reshape(self.ncell, 6), f.reshape(self.natom, 3), rho

    def tune(self, reps=3):
        b, g, ms = C.c_int(), C.c_int(), C.c_double()
        _bcheck(self.L.mdk_batch_tune(self.h, reps, C.byref(b), C.byref(g), C.byref(ms)))
        return b.value, g.value, ms.value

    def time(self, reps=10) -> float:
        ms = C.c_double()
        _bcheck(self.L.mdk_batch_time(self.h, reps, C.byref(ms)))
        return ms.value

    def set_targets(self, e_target, f_target):
        """Loss targets: per-cell energies (eV) and Cartesian forces (natom, 3)."""
        self._e_t = np.ascontiguousarray(e_target, dtype=np.float64).reshape(self.ncell)
        self._f_t = np.ascontiguousarray(f_target, dtype=np.float64).reshape(3 * self.natom)
        _bcheck(self.L.mdk_batch_set_targets(self.h, _dp(self._e_t), _dp(self._f_t)))

    def loss_grad(self):
        """(loss, d loss / d theta) with theta = (a_k, r_k, A_k, R_k)."""
        loss = C.c_double()
        g = np.empty(self.ngrad)
        _bcheck(self.L.mdk_batch_loss_grad(self.h, C.byref(loss), _dp(g)))
        return loss.value, g

    def loss_grad_noalloc(self, loss_ptr: int, grad_ptr: int):
        """loss_grad into device memory (1 and ngrad doubles at the given device
        addresses) with no allocation; asynchronous: synchronize() before reading."""
        _bcheck(self.L.mdk_batch_loss_grad_noalloc(self.h, C.c_void_p(loss_ptr), C.c_void_p(grad_ptr)))

    def vjp(self, dE, dF):
        """sum_c dE_c dE_c/dtheta + sum_i dF_i . dF_i/dtheta at the last forward evaluation."""
        de = np.ascontiguousarray(dE, dtype=np.float64).reshape(self.ncell)
        df = np.ascontiguousarray(dF, dtype=np.float64).reshape(3 * self.natom)
        g = np.empty(self.ngrad)
        _bcheck(self.L.mdk_batch_vjp(self.h, _dp(de), _dp(df), _dp(g)))
        return g

    def time_loss_grad(self, reps=10) -> float:
        ms = C.c_double()
        _bcheck(self.L.mdk_batch_time_loss_grad(self.h, reps, C.byref(ms)))
        return ms.value

    def info(self):
        n, b, g, p = C.c_int(), C.c_int(), C.c_int(), C.c_long()
        _bcheck(self.L.mdk_batch_info(self.h, C.byref(n), C.byref(b), C.byref(g), C.byref(p)))
        return {"natom": n.value, "block": b.value, "wg_per_cu": g.value, "pairs": p.value}

    def close(self):
        if getattr(self, "h", None):
            self.L.mdk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def virial_to_stress(w6: np.ndarray, b0: np.ndarray) -> np.ndarray:
    """stress = -W/V (the reference's efs.out sigma, src/dynamics_m.F90:610-615)."""
    W = np.array([[w6[0], w6[3], w6[4]], [w6[3], w6[1], w6[5]], [w6[4], w6[5], w6[2]]])
    return -W / abs(np.linalg.det(b0))
