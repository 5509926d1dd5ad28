"""Calibration workload of the 4D-Var layer: Manning zones, an inflow
hydrograph and gauges on the SP mesh at scale.

The reference calibrates porosity per cell (assim/swSpE_checkpoint.cpp) and
keeps friction and inflow constant; this workload instead calibrates 68
controls -- the Manning roughness n of 8 zones (k-means clusters of the cell
centroids) and 60 knots of the west inflow hydrograph -- from 20 depth gauges
sampled every 30 s, through the same differentiable forward and adjoint.
Gauges sit at seeded points located in their cell by a point-in-triangle
test (the role of getCellFromCoordinates, src/common/geometry/swMesh.cpp:695-759).
"""
import numpy as np

from . import _native as N
from .cases import scale_case


def kmeans_zones(x, y, k, seed, iters=25, fit_max=50000):
    """Lloyd's k-means of the cell centroids (fitted on a seeded subsample of
    at most fit_max cells, then every cell assigned to its nearest centre);
    zones renumbered by centre x."""
    rng = np.random.default_rng(seed)
    P = np.stack([x, y], axis=1)
    S = P[rng.choice(len(P), min(len(P), fit_max), replace=False)]
    C = S[rng.choice(len(S), k, replace=False)].copy()
    for _ in range(iters):
        lab = np.argmin(((S[:, None, :] - C[None, :, :]) ** 2).sum(axis=2), axis=1)
        for j in range(k):
            sel = lab == j
            if np.any(sel):
                C[j] = S[sel].mean(axis=0)
    lab = np.argmin(((P[:, None, :] - C[None, :, :]) ** 2).sum(axis=2), axis=1)
    order = np.argsort(C[:, 0])
    rank = np.empty(k, dtype=np.int64)
    rank[order] = np.arange(k)
    return rank[lab]


def locate_cells(geom, px, py):
    """Cell containing each point (-1 if outside): sign test against the
    three edges of every triangle, nearest centroid first."""
    nx_, ny_ = geom.array("node_x"), geom.array("node_y")
    tri = geom.array("cell_tri").reshape(-1, 3)
    cx, cy = nx_[tri].mean(axis=1), ny_[tri].mean(axis=1)
    out = []
    for x, y in zip(np.atleast_1d(px), np.atleast_1d(py)):
        cand = np.argsort((cx - x) ** 2 + (cy - y) ** 2)[:16]
        found = -1
        for c in cand:
            a, b, d = tri[c]
            s = []
            for (i, j) in ((a, b), (b, d), (d, a)):
                s.append((nx_[j] - nx_[i]) * (y - ny_[i]) - (ny_[j] - ny_[i]) * (x - nx_[i]))
            if (min(s) >= 0.0) or (max(s) <= 0.0):
                found = int(c)
                break
        out.append(found)
    return np.array(out, dtype=np.int64)


class Calibration:
    """The workload: mesh + truth / default controls + gauges.

    Controls: n_zone[8] (Manning n per zone; the _E forward applies the
    Strickler law, so K = 1/n per cell) and Q[60] (m3/s at the knots
    t_knots, turned into unit discharge on the west edges)."""

    def __init__(self, px=707, seed=1234, n_zones=8, n_knots=60, n_gauges=20, obs_dt=30.0,
                 window=90.0, q_peak=500.0, t_event=3600.0):
        self.window = float(window)
        self.t_knots = np.linspace(0.0, self.window, n_knots)
        # default hydrograph: the rising limb of the event's triangular
        # hydrograph (0 -> q_peak over t_event / 2) on the knots
        base = q_peak * np.minimum(self.t_knots / (0.5 * t_event), 1.0) + 5.0
        self.case, info = scale_case(None, px=px, py=px, L=5000.0, seed=seed, t_max=self.window,
                                     scheme="godunov", friction="strickler",
                                     manning=1.0 / 0.03, q_peak=q_peak,
                                     hydrograph=(self.t_knots, base))
        self.info = info
        g = info["geometry"]
        a = self.case.a
        rng = np.random.default_rng(seed + 1)
        self.n_zones = n_zones
        self.zone = kmeans_zones(a["cell_x"], a["cell_y"], n_zones, seed + 2)
        self.n_true = rng.uniform(0.015, 0.06, n_zones)
        self.n_default = np.full(n_zones, 0.03)
        self.Q_default = base
        # truth: the same limb with a seeded +-30 % random-walk wiggle
        wig = np.cumsum(rng.standard_normal(n_knots))
        wig = (wig - wig.mean()) / (np.abs(wig - wig.mean()).max() + 1e-12)
        self.Q_true = base * (1.0 + 0.3 * wig)
        self.west_len = info["west_len"]
        # initial state: a free surface sloping down to the east at half the
        # bed slope, at rest -- water moves everywhere from the start, so the
        # roughness of every zone shapes the gauge series; dry where the
        # bumps emerge and near the west edge, where the inflow wets the bed
        zs = info["z0"] + 0.5e-3 * (5000.0 - a["cell_x"])
        a["h0"] = np.maximum(zs - a["cell_zb"], 0.0)
        a["qx0"] = np.zeros(self.case.n_cell)
        a["qy0"] = np.zeros(self.case.n_cell)
        # gauges: a third at seeded points near the inflow (x < 300 m), the
        # rest at seeded wet points anywhere
        h0 = a["h0"]
        cells = []
        n_west = n_gauges * 2 // 5
        while len(cells) < n_gauges:
            west = len(cells) < n_west
            x = rng.uniform(10.0, 300.0) if west else rng.uniform(0.0, 5000.0)
            y = rng.uniform(0.0, 5000.0)
            c = int(locate_cells(g, [x], [y])[0])
            if c >= 0 and (west or h0[c] > 0.2) and c not in cells:
                cells.append(c)
        self.gauge_cells = np.array(cells, dtype=np.int64)
        self.obs_t = np.arange(obs_dt, self.window + 1e-9, obs_dt)
        self.series = 1  # west inflow series (boundary code 2)

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
