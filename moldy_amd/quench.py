"""Static relaxation (quench): minimum of the enthalpy H = pe + press vol over
the fractional positions and, for ivol < 1 (or 4: the zz element only), the
box, as MOLDY's quench does (src/quench_m.F90:100-322).

The objective and its gradient are the reference's derivs
(src/quench_m.F90:214-322): for atom i, dH/ds_i = -tg f_i with tg = B0^T B0
and f_i the fractional force (so -B0^T of the Cartesian force); for the box,
dH/dB0(i,j) = -tp(i,j) + press tc(i,j); the variables are stored as the
reference stores them (positions atom by atom, then B0 row by row).  Energy,
forces, tp and tc come from one device force evaluation per point.

The minimiser is this project's own Polak-Ribiere conjugate gradient (with
the PR+ reset and a restart every n steps) and a line search that brackets
the minimum along the direction and refines it by safeguarded secant steps on
the directional derivative, stopping at |g.d| <= 0.1 |g0.d| (strong Wolfe
curvature).  The reference uses the Harwell routine VA14CD, whose licence keeps
it out of GPL code (src/quench_m.F90:325-335); as there, the first step is
sized for an expected decrease dfn and the run ends when an iteration lowers
H by less than acc (default 5e-7 nm, src/quench_m.F90:134-136).
"""
from __future__ import annotations

import numpy as np


class _Objective:
    def __init__(self, eng, ivol: int, press: float):
        self.eng, self.ivol, self.press = eng, ivol, press
        self.n = eng.n
        self.nevals = 0
        if ivol < 1:
            self.bmask = np.ones(9, dtype=bool)
        elif ivol == 4:
            self.bmask = np.zeros(9, dtype=bool)
            self.bmask[8] = True
        else:
            self.bmask = np.zeros(9, dtype=bool)

    def __call__(self, x: np.ndarray):
        n, eng = self.n, self.eng
        s = x[:3 * n].reshape(n, 3)
        b0 = x[3 * n:].reshape(3, 3)            # row by row: b0(i,j) at 3(i-1)+(j-1)
        for k, nm in enumerate(("x0", "y0", "z0")):
            eng.upload(nm, s[:, k])
        eng.set_tensor("b0", b0)
        eng.force()
        self.nevals += 1
        sc = eng.scalars()
        pe, vol = sc["pe"], sc["vol"]
        f = np.stack([eng.array("fx"), eng.array("fy"), eng.array("fz")], axis=1)
        tg = b0.T @ b0
        g = np.empty_like(x)
        g[:3 * n] = (-(f @ tg.T)).ravel()
        gb = -eng.tensor("tp") + self.press * eng.tensor("tc")
        g[3 * n:] = np.where(self.bmask, gb.ravel(), 0.0)
        return pe + self.press * vol, g, pe, vol


def _line_search(obj, x, fx, gx, d, step0, max_evals=40):
    """Minimum of phi(a) = H(x + a d) for a > 0: bracket with growing steps,
    then safeguarded secant steps on phi'(a) = g(x + a d).d."""
    d0 = float(gx @ d)
    lo, flo, dlo = 0.0, fx, d0
    a = step0
    hi = None
    best = (fx, x, gx, 0.0)
    for _ in range(max_evals):
        xa = x + a * d
        fa, ga, *_ = obj(xa)
        da = float(ga @ d)
        if fa < best[0]:
            best = (fa, xa, ga, a)
        if abs(da) <= 0.1 * abs(d0) and fa <= fx:
            return fa, xa, ga, a
        if da > 0.0 or fa > flo:
            hi, fhi, dhi = a, fa, da
        else:
            lo, flo, dlo = a, fa, da
        if hi is None:
            a = 2.0 * a if da < 0.0 else a
            continue
        # secant on the derivative inside [lo, hi], kept away from the ends
        if dhi > 0.0 and dlo < 0.0:
            t = lo + (hi - lo) * (-dlo) / (dhi - dlo)
        else:
            t = 0.5 * (lo + hi)
        w = hi - lo
        a = min(max(t, lo + 0.1 * w), hi - 0.1 * w)
    return best


def quench(eng, ivol: int, press: float, acc: float | None = None, dfn: float = 0.1,
           max_iter: int = 20000) -> dict:
    """Relaxes the engine's state in place; returns the final enthalpy,
    energy, volume, iteration and evaluation counts.  press in model units."""
    n = eng.n
    acc = 5.0e-7 * n if acc is None else acc
    obj = _Objective(eng, ivol, press)
    s = np.stack([eng.array("x0"), eng.array("y0"), eng.array("z0")], axis=1)
    x = np.concatenate([s.ravel(), eng.tensor("b0").ravel()])
    fx, g, pe, vol = obj(x)
    d = -g
    step = 2.0 * dfn / max(float(g @ g), 1e-300)     # expected first decrease dfn
    small = 0
    it = 0
    for it in range(1, max_iter + 1):
        gd = float(g @ d)
        if gd >= 0.0:                               # not a descent direction: restart
            d = -g
            gd = -float(g @ g)
        f_new, x_new, g_new, a = _line_search(obj, x, fx, g, d, step)
        if a == 0.0:                                # no decrease along d
            if np.array_equal(d, -g):
                break
            d = -g
            continue
        dec = fx - f_new
        beta = max(0.0, float(g_new @ (g_new - g)) / max(float(g @ g), 1e-300))   # PR+
        if it % max(3 * n + 9, 50) == 0:
            beta = 0.0
        d_new = -g_new + beta * d
        gd_new = float(g_new @ d_new)
        # next first step from the last one scaled by the directional derivatives
        step = a * gd / gd_new if gd_new < 0.0 else a
        x, fx, g, d = x_new, f_new, g_new, d_new
        small = small + 1 if dec < acc else 0
        if small >= 2:
            break
    fx, g, pe, vol = obj(x)
    return dict(h=fx, pe=pe, vol=vol, iterations=it, evaluations=obj.nevals, gmax=float(np.max(np.abs(g))))
