"""Calibration workload of the 4D-Var layer: Manning zones, an inflow
hydrograph and gauges on the SP mesh at scale, and the 4D-Var that
calibrates them.

The reference calibrates porosity per cell (assim/swSpE_checkpoint.cpp,
minimised by L-BFGS-B in assim/swSpE_minimize.cpp:267-1073) and keeps
friction and inflow constant; this workload calibrates 68 controls -- the
Manning roughness n of 8 zones and 60 knots of the west inflow hydrograph --
from 20 depth gauges sampled every 30 s, through the same differentiable
forward and the GPU adjoint.  Gauges sit at seeded points located in their
cell by a point-in-triangle test (the role of getCellFromCoordinates,
src/common/geometry/swMesh.cpp:695-759).

Every quantity that defines the physical problem (bed, porosity field, zone
centres, gauge points, start level, hydrographs) is a function of position
and seed only, so meshes of different resolution of the same domain carry
the same problem: the batched ensemble runs coarse meshes of it, the 4D-Var
the fine one.
"""
import functools
import time

import numpy as np

from . import _native as N
from .cases import scale_case

L_DOMAIN = 5000.0


@functools.lru_cache(maxsize=16)
def zone_centres(k, seed, L=L_DOMAIN, iters=25, n_fit=50000):
    """Centres of k Manning zones: Lloyd's k-means of seeded uniform points of
    the L x L domain, ordered by x."""
    rng = np.random.default_rng(seed)
    S = rng.uniform(0.0, L, (n_fit, 2))
    C = S[rng.choice(len(S), k, replace=False)].copy()
    for _ in range(iters):
        lab = np.argmin(((S[:, None, :] - C[None, :, :]) ** 2).sum(axis=2), axis=1)
        for j in range(k):
            sel = lab == j
            if np.any(sel):
                C[j] = S[sel].mean(axis=0)
    return C[np.argsort(C[:, 0])]


def kmeans_zones(x, y, k, seed, L=L_DOMAIN):
    """Zone of every cell: the nearest of the k zone centres."""
    C = zone_centres(k, seed, L)
    P = np.stack([x, y], axis=1)
    return np.argmin(((P[:, None, :] - C[None, :, :]) ** 2).sum(axis=2), axis=1).astype(np.int64)


def locate_cells(geom, px, py):
    """Cell containing each point (-1 if outside): locate_gauges
    (assim/swSpE_gauges.cpp:41-77) through the reference's winding-angle test
    (swMesh::getCellFromCoordinates, swMesh.cpp:695-759), any polygon mesh."""
    return geom.locate(px, py)


def bed(bumps, x, y, L=L_DOMAIN):
    """Bed elevation of the scale case at points (x, y): slope 1e-3 down to
    the east plus the Gaussian bumps."""
    z = 1e-3 * (L - np.asarray(x, dtype=np.float64))
    for cx, cy, amp, sig in bumps:
        z = z + amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * sig * sig))
    return z


class Calibration:
    """The workload on one mesh: mesh + truth / default controls + gauges.

    Controls: n_zone[8] (Manning n per zone; the _E forward applies the
    Strickler law, so K = 1/n per cell) and Q[60] (m3/s at the knots
    t_knots, turned into unit discharge on the west edges)."""

    def __init__(self, px=707, seed=1234, n_zones=8, n_knots=60, n_gauges=20, obs_dt=30.0,
                 window=90.0, q_peak=500.0, t_event=3600.0, truth_seed=None, level_quantile=0.4,
                 scheme="godunov"):
        self.window = float(window)
        self.t_knots = np.linspace(0.0, self.window, n_knots)
        # default hydrograph: the rising limb of the event's triangular
        # hydrograph (0 -> q_peak over t_event / 2) on the knots
        base = q_peak * np.minimum(self.t_knots / (0.5 * t_event), 1.0) + 5.0
        self.case, info = scale_case(None, px=px, py=px, L=L_DOMAIN, seed=seed, t_max=self.window,
                                     scheme=scheme, friction="strickler",
                                     manning=1.0 / 0.03, q_peak=q_peak,
                                     hydrograph=(self.t_knots, base))
        self.info = info
        g = info["geometry"]
        a = self.case.a
        self.n_zones = n_zones
        self.zone = kmeans_zones(a["cell_x"], a["cell_y"], n_zones, seed + 2)
        self.n_default = np.full(n_zones, 0.03)
        self.Q_default = base
        self.n_true, self.Q_true = self.sample_controls(seed + 1 if truth_seed is None else truth_seed)
        self.west_len = info["west_len"]
        # reference level from the analytic bed (mesh independent); initial
        # state: a free surface sloping down to the east at half the bed
        # slope, at rest -- water moves everywhere from the start, so the
        # roughness of every zone shapes the gauge series; dry where the
        # bumps emerge and near the west edge, where the inflow wets the bed
        bumps = info["bumps"]
        gx, gy = np.meshgrid(np.linspace(0, L_DOMAIN, 256), np.linspace(0, L_DOMAIN, 256))
        self.z0 = float(np.quantile(bed(bumps, gx.ravel(), gy.ravel()), level_quantile))
        zs = self.z0 + 0.5e-3 * (L_DOMAIN - a["cell_x"])
        a["h0"] = np.maximum(zs - a["cell_zb"], 0.0)
        a["qx0"] = np.zeros(self.case.n_cell)
        a["qy0"] = np.zeros(self.case.n_cell)
        # gauge points: 2/5 near the inflow (x < 300 m), the rest where the
        # analytic start depth exceeds 0.2 m
        rng = np.random.default_rng(seed + 3)
        pts = []
        n_west = n_gauges * 2 // 5
        while len(pts) < n_gauges:
            west = len(pts) < n_west
            x = rng.uniform(10.0, 300.0) if west else rng.uniform(0.0, L_DOMAIN)
            y = rng.uniform(0.0, L_DOMAIN)
            if west or (self.z0 + 0.5e-3 * (L_DOMAIN - x) - bed(bumps, x, y)) > 0.2:
                pts.append((x, y))
        self.gauge_xy = np.array(pts)
        self.gauge_cells = locate_cells(g, self.gauge_xy[:, 0], self.gauge_xy[:, 1])
        self.obs_t = np.arange(obs_dt, self.window + 1e-9, obs_dt)
        self.series = 1  # west inflow series (boundary code 2)

    def sample_controls(self, seed):
        """Seeded controls of the problem's prior: n per zone uniform in
        [0.015, 0.06]; the default hydrograph times a +-30 % random walk."""
        rng = np.random.default_rng(seed)
        n = rng.uniform(0.015, 0.06, self.n_zones)
        wig = np.cumsum(rng.standard_normal(len(self.t_knots)))
        wig = (wig - wig.mean()) / (np.abs(wig - wig.mean()).max() + 1e-12)
        return n, self.Q_default * (1.0 + 0.3 * wig)

    def K(self, n_zone):
        return 1.0 / np.asarray(n_zone, dtype=np.float64)[self.zone]

    def bc_values(self, Q):
        """[n_series, n_times] table of the two q series (walls, west inflow)."""
        v = np.zeros((2, len(self.t_knots)))
        v[1] = np.asarray(Q, dtype=np.float64) / self.west_len
        return v

    def apply_hydrograph(self, solver, Q):
        solver.set_bc_series(self.t_knots, self.bc_values(Q),
                             types=np.array([N.BC_Q, N.BC_Q], dtype=np.int32))

    def grad_controls(self, G, n_zone, t_steps):
        """dJ/dn_zone and dJ/dQ_knot from the per-cell / per-step gradients."""
        from .assim import interp_weights
        n_zone = np.asarray(n_zone, dtype=np.float64)
        gK = np.zeros(self.n_zones)
        np.add.at(gK, self.zone, G["K1"] + G["K2"])
        gn = gK * (-1.0 / n_zone ** 2)
        W = interp_weights(self.t_knots, t_steps)
        gQ = (W.T @ G["bc"][:, self.series]) / self.west_len
        return gn, gQ

    # -- twin experiment and 4D-Var -------------------------------------------
    def observe(self, solver, n_zone, Q, max_steps=200000):
        """Gauge series [n_obs, n_gauges] of a run at the given controls."""
        from .assim import Assim
        self.apply_hydrograph(solver, Q)
        K = self.K(n_zone)
        A = Assim(solver, self.obs_t, gauge_cells=self.gauge_cells,
                  gauge_h=np.zeros((len(self.obs_t), len(self.gauge_cells))), max_steps=max_steps)
        A.record(K1=K, K2=K)
        self._check_window(A.get_record(), max_steps)
        return A.gauge_series()

    def _check_window(self, rec, max_steps):
        if rec["t"][-1] < self.obs_t[-1] - 1e-9 * max(1.0, self.obs_t[-1]):
            raise RuntimeError(f"the run stopped at t = {rec['t'][-1]:.3f} s, before the last "
                               f"observation at {self.obs_t[-1]:.3f} s: raise max_steps "
                               f"(now {max_steps})")

    def fourdvar(self, solver, obs, n0, Q0, J_tol=0.0, max_iter=60, max_steps=200000, m=6,
                 n_bounds=(0.01, 0.1), Q_bounds=(0.0, 5000.0), max_ls=30, time_limit=None,
                 n_tiles=1, ls_cost="auto", adj_cut=0.0, log=None):
        """Projected L-BFGS (memory m) on the 68 controls from (n0, Q0)
        against the gauge series `obs`, organised like the reference's
        minimiser (assim/swSpE_minimize.cpp:398-511): a gradient evaluation
        runs the dry run at the controls (fixing the time-step schedule) and
        the adjoint; the line-search probes evaluate the cost on that frozen
        schedule (replay only, ls_cost "frozen").  With n_tiles > 1 the
        gradient is that of the tiled cost (each tile restarts from the dry
        run's state at its top), which drifts from the true misfit away from
        the recorded trajectory; the probes then run a dry run at the trial
        controls and take the true misfit of its gauge series (ls_cost
        "record", the default for tiles).  adj_cut > 0 drops the adjoint
        of thin films (sw2dg_assim_set_adjoint_cut); where the line search
        then finds no decrease (the cut gradient need not be a descent
        direction), the gradient at that point is evaluated again with the
        exact adjoint and the search repeated along it before giving up;
        the cut applies again after the step.  Controls
        are scaled to O(1)
        (n / 0.03, Q / max Q_default) and kept in their bounds by projection.
        Stops when J <= J_tol, after max_iter iterations or time_limit
        seconds.  Returns (n, Q, trace) with trace = [(wall s, J, kind)] for
        every evaluation, kind 'grad' or 'ls' (each also passed to `log`, if
        given, as it happens)."""
        from .assim import Assim, interp_weights
        A = Assim(solver, self.obs_t, gauge_cells=self.gauge_cells, gauge_h=obs,
                  max_steps=max_steps, n_tiles=n_tiles, adj_cut=adj_cut)
        nz = self.n_zones
        sn, sq = 0.03, max(float(np.max(self.Q_default)), 1.0)
        lo = np.concatenate([np.full(nz, n_bounds[0] / sn), np.full(len(Q0), Q_bounds[0] / sq)])
        hi = np.concatenate([np.full(nz, n_bounds[1] / sn), np.full(len(Q0), Q_bounds[1] / sq)])
        trace = []
        t0 = time.perf_counter()
        state = {}

        def note(J, kind):
            trace.append((time.perf_counter() - t0, J, kind))
            if log:
                log(f"4D-Var {kind}: {trace[-1][0]:.1f} s, J = {J:.4e}")

        def controls(z):
            return z[:nz] * sn, z[nz:] * sq

        def grad(z):
            n, Q = controls(z)
            self.apply_hydrograph(solver, Q)
            K = self.K(n)
            G = A.gradient(K1=K, K2=K)
            rec = A.get_record()
            self._check_window(rec, max_steps)
            gn, gQ = self.grad_controls(G, n, rec["t"][:-1])
            state["bc"] = rec["bc"].copy()
            state["W"] = interp_weights(self.t_knots, rec["t"][:-1])
            note(G["J"], "grad")
            return G["J"], np.concatenate([gn * sn, gQ * sq])

        if ls_cost == "auto":
            ls_cost = "frozen" if n_tiles == 1 else "record"
        obs_arr = np.asarray(obs, dtype=np.float64)

        def cost(z):
            n, Q = controls(z)
            if ls_cost == "record":  # true misfit of a dry run at z
                self.apply_hydrograph(solver, Q)
                K = self.K(n)
                A.record(K1=K, K2=K)
                self._check_window(A.get_record(), max_steps)
                J = float(np.sum((A.gauge_series() - obs_arr) ** 2))
                note(J, "ls")
                return J
            # frozen schedule of the last gradient evaluation
            bc = state["bc"].copy()
            bc[:, self.series] = state["W"] @ Q / self.west_len
            A.set_bc_seq(bc)
            K = self.K(n)
            J = A.cost(K1=K, K2=K)
            note(J, "ls")
            return J

        def search(z, J, g, d, step):
            """Backtracking (Armijo) along d from `step` (None: a first step
            of 0.1 in the largest control); the accepted point, or None."""
            if step is None:
                step = min(1.0, 0.1 / max(np.max(np.abs(d)), 1e-300))
            for _ in range(max_ls):
                zn = np.clip(z + step * d, lo, hi)
                if cost(zn) <= J + 1e-4 * np.dot(g, zn - z):
                    return zn
                step *= 0.5
            return None

        z = np.clip(np.concatenate([np.asarray(n0) / sn, np.asarray(Q0) / sq]), lo, hi)
        J, g = grad(z)
        S, Y = [], []
        for it in range(max_iter):
            if J <= J_tol or (time_limit and time.perf_counter() - t0 > time_limit):
                break
            free = ~(((z <= lo) & (g > 0)) | ((z >= hi) & (g < 0)))
            q = np.where(free, g, 0.0)
            alphas = []
            for s_, y_ in reversed(list(zip(S, Y))):
                a = np.dot(s_, q) / np.dot(y_, s_)
                alphas.append(a)
                q = q - a * y_
            if S:
                q = q * (np.dot(S[-1], Y[-1]) / np.dot(Y[-1], Y[-1]))
            for (s_, y_), a in zip(zip(S, Y), reversed(alphas)):
                b = np.dot(y_, q) / np.dot(y_, s_)
                q = q + s_ * (a - b)
            d = -np.where(free, q, 0.0)
            if np.dot(d, g) >= 0.0:  # not a descent direction: restart from steepest descent
                S, Y = [], []
                d = -np.where(free, g, 0.0)
            # the quasi-Newton direction, then steepest descent (with a cut
            # adjoint: along the exact gradient, evaluated here)
            zn = search(z, J, g, d, 1.0 if S else None)
            if zn is None and S and A.adj_cut == 0.0:
                S, Y = [], []
                zn = search(z, J, g, -np.where(free, g, 0.0), None)
            if zn is None and A.adj_cut > 0.0:
                cut = A.adj_cut
                A.set_adjoint_cut(0.0)
                J, g = grad(z)
                A.set_adjoint_cut(cut)
                S, Y = [], []
                free = ~(((z <= lo) & (g > 0)) | ((z >= hi) & (g < 0)))
                zn = search(z, J, g, -np.where(free, g, 0.0), None)
            if zn is None:
                break
            Jn, gn = grad(zn)
            sk, yk = zn - z, gn - g
            if np.dot(sk, yk) > 1e-12 * np.dot(sk, sk):
                S.append(sk)
                Y.append(yk)
                if len(S) > m:
                    S.pop(0)
                    Y.pop(0)
            z, J, g = zn, Jn, gn
        n, Q = controls(z)
        return n, Q, trace
