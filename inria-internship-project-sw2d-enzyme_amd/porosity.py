"""Porosity calibration problem and its minimiser: the host mirror of the
reference's PorosityProblem (assim/swSpE_minimize.cpp:267-511) over the GPU
gradient (Assim).

The optimisation variable theta is the control m on the free cells (or one
value per zone of equal background porosity with zonation); fixed cells keep
the background phi_b.  f(theta) = J_data(m) on the frozen trajectory of the
last gradient + ridge + L1 + spatial (H1) + phi_ref penalties, all on
phi = phi_correct(m) (the change of variable of the Assim, none / clamp /
tanh / normalize).  As in the reference, the J that grad() returns leaves
the spatial term out (swSpE_minimize.cpp:487) while its gradient includes
it; func() includes it.

``minimize`` is a projected L-BFGS (memory 6, at most 30 line-search
probes, the reference's settings for its functional case) over box bounds.
"""
import time

import numpy as np

from .assim import phi_transform_eval


class PorosityProblem:
    def __init__(self, assim, phi_b, free_idx=None, alpha=0.0, alpha_l1=0.0, alpha_spatial=0.0,
                 alpha_phi_ref=0.0, phi_ref=None, zonation=False, edges=None,
                 phi_transform=None, K1=None, K2=None):
        """assim: an Assim built with the same phi_transform; edges = (L, R)
        internal-edge cell indices (reference order) for the spatial term."""
        self.A = assim
        self.phi_b = np.asarray(phi_b, dtype=np.float64)
        self.n = len(self.phi_b)
        self.free = (np.arange(self.n) if free_idx is None
                     else np.asarray(free_idx, dtype=np.int64))
        self.alpha, self.alpha_l1 = float(alpha), float(alpha_l1)
        self.alpha_spatial, self.alpha_phi_ref = float(alpha_spatial), float(alpha_phi_ref)
        self.phi_ref = None if phi_ref is None else np.asarray(phi_ref, dtype=np.float64)
        self.edges = edges
        self.mode, self.lam = (None, 1.0) if phi_transform is None else (tuple(phi_transform) + (1.0,))[:2]
        self.K1, self.K2 = K1, K2
        self.zone_of_free = None
        if zonation:  # cells of equal background (quantised to 1e-6) move together
            keys = np.round(self.phi_b[self.free] * 1e6).astype(np.int64)
            _, first, inv = np.unique(keys, return_index=True, return_inverse=True)
            # zones numbered in order of first appearance, as the reference's map insertion
            order = np.argsort(first)
            rank = np.empty_like(order)
            rank[order] = np.arange(len(order))
            self.zone_of_free = rank[inv]
            self.zone_value = self.phi_b[self.free][np.sort(first)]
        self.grad_calls = 0

    # -- variables ---------------------------------------------------------------
    def size(self):
        return int(self.zone_of_free.max() + 1) if self.zone_of_free is not None else len(self.free)

    def theta0(self):
        """The background as the starting variable (phi_b on the free cells / zones)."""
        if self.zone_of_free is not None:
            return self.zone_value.copy()
        return self.phi_b[self.free].copy()

    def m_of(self, theta):
        m = self.phi_b.copy()
        theta = np.asarray(theta, dtype=np.float64)
        m[self.free] = theta[self.zone_of_free] if self.zone_of_free is not None else theta
        return m

    def _phi(self, m):
        lo = hi = 0.0
        if self.mode == "normalize":  # bounds of the dry run (update_phi_norm_bounds)
            lo, hi = float(m[self.free].min()), float(m[self.free].max())
        return phi_transform_eval(self.mode, m, lam=self.lam, lo=lo, hi=hi)

    # -- penalties (swSpE_minimize.cpp:345-396) ----------------------------------
    def penalties(self, m, phi=None):
        phi = self._phi(m)[0] if phi is None else phi
        f = self.free
        out = {"ridge": 0.0, "l1": 0.0, "spatial": 0.0, "phi_ref": 0.0}
        if self.alpha > 0.0:
            out["ridge"] = 0.5 * self.alpha * float(np.sum((phi[f] - self.phi_b[f]) ** 2))
        if self.alpha_l1 > 0.0:
            out["l1"] = self.alpha_l1 * float(np.sum(np.abs(phi[f] - self.phi_b[f])))
        if self.alpha_spatial > 0.0 and self.edges is not None:
            L, R = self.edges
            out["spatial"] = 0.5 * self.alpha_spatial * float(np.sum((phi[L] - phi[R]) ** 2))
        if self.alpha_phi_ref > 0.0 and self.phi_ref is not None:
            out["phi_ref"] = 0.5 * self.alpha_phi_ref * float(np.sum((phi[f] - self.phi_ref[f]) ** 2))
        return out

    def _ctl(self):
        kw = {}
        if self.K1 is not None:
            kw["K1"] = self.K1
        if self.K2 is not None:
            kw["K2"] = self.K2
        return kw

    # -- f and its gradient (swSpE_minimize.cpp:398-511) --------------------------
    def func(self, theta):
        m = self.m_of(theta)
        J = self.A.cost(phi=m, **self._ctl())
        return J + sum(self.penalties(m).values())

    def grad(self, theta):
        m = self.m_of(theta)
        G = self.A.gradient(phi=m, **self._ctl())
        J_data = self.A.cost(phi=m, **self._ctl())
        phi, dphi = self._phi(m)
        f = self.free
        g = G["phi"][f].copy()  # dJ_data/dm (the change of variable is in the library)
        g += self.alpha * (phi[f] - self.phi_b[f]) * dphi[f]
        if self.alpha_l1 > 0.0:
            diff = phi[f] - self.phi_b[f]
            g += self.alpha_l1 * np.where(diff > 1e-12, 1.0, np.where(diff < -1e-12, -1.0, 0.0)) * dphi[f]
        if self.alpha_spatial > 0.0 and self.edges is not None:
            L, R = self.edges
            d = self.alpha_spatial * (phi[L] - phi[R])
            gs = np.zeros(self.n)
            np.add.at(gs, L, d)
            np.add.at(gs, R, -d)
            g += gs[f] * dphi[f]
        if self.alpha_phi_ref > 0.0 and self.phi_ref is not None:
            g += self.alpha_phi_ref * (phi[f] - self.phi_ref[f]) * dphi[f]
        if self.zone_of_free is not None:
            gz = np.zeros(self.size())
            np.add.at(gz, self.zone_of_free, g)
            g = gz
        p = self.penalties(m, phi)
        self.grad_calls += 1
        # the reference's grad() J: data + ridge + L1 + phi_ref (no spatial term)
        return J_data + p["ridge"] + p["l1"] + p["phi_ref"], g


def minimize(fg, f, x0, lo=None, hi=None, m=6, max_iter=100, max_ls=30, eps=1e-5, time_limit=None):
    """Projected L-BFGS on box bounds: fg(x) -> (value, gradient) anchors the
    trajectory, f(x) evaluates line-search probes.  Stops when the projected
    gradient's max norm falls below eps, or at max_iter / time_limit.
    Returns (x, value, trace [(wall s, value, 'grad' | 'ls')])."""
    x = np.asarray(x0, dtype=np.float64).copy()
    lo = np.full_like(x, -np.inf) if lo is None else np.broadcast_to(np.asarray(lo, float), x.shape)
    hi = np.full_like(x, np.inf) if hi is None else np.broadcast_to(np.asarray(hi, float), x.shape)
    x = np.clip(x, lo, hi)
    t0 = time.perf_counter()
    trace = []
    J, g = fg(x)
    trace.append((time.perf_counter() - t0, J, "grad"))
    S, Y = [], []
    for _ in range(max_iter):
        free = ~(((x <= lo) & (g > 0)) | ((x >= hi) & (g < 0)))
        if np.max(np.abs(np.where(free, g, 0.0)), initial=0.0) <= eps:
            break
        if time_limit and time.perf_counter() - t0 > time_limit:
            break
        q = np.where(free, g, 0.0)
        alphas = []
        for s_, y_ in reversed(list(zip(S, Y))):
            a = np.dot(s_, q) / np.dot(y_, s_)
            alphas.append(a)
            q = q - a * y_
        if S:
            q = q * (np.dot(S[-1], Y[-1]) / np.dot(Y[-1], Y[-1]))
        for (s_, y_), a in zip(zip(S, Y), reversed(alphas)):
            q = q + s_ * (a - np.dot(y_, q) / np.dot(y_, s_))
        d = -np.where(free, q, 0.0)
        if np.dot(d, g) >= 0.0:
            S, Y = [], []
            d = -np.where(free, g, 0.0)
        step = 1.0 if S else min(1.0, 0.1 / max(np.max(np.abs(d)), 1e-300))
        accepted = False
        for _ in range(max_ls):
            xn = np.clip(x + step * d, lo, hi)
            Jn = f(xn)
            trace.append((time.perf_counter() - t0, Jn, "ls"))
            if Jn <= J + 1e-4 * np.dot(g, xn - x):
                accepted = True
                break
            step *= 0.5
        if not accepted:
            break
        Jn, gn = fg(xn)
        trace.append((time.perf_counter() - t0, Jn, "grad"))
        sk, yk = xn - x, gn - g
        if np.dot(sk, yk) > 1e-12 * np.dot(sk, sk):
            S.append(sk)
            Y.append(yk)
            if len(S) > m:
                S.pop(0)
                Y.pop(0)
        x, J, g = xn, Jn, gn
    return x, J, trace
