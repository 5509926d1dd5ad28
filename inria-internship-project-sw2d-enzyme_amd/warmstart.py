"""Learned warm start of the calibration: ensemble -> MLP -> 4D-Var.

1. ``training_set``: gauge series of many runs of the calibration problem
   (Calibration: 8 Manning zones, 60-knot inflow hydrograph, 20 gauges every
   30 s) at controls drawn from the problem's prior, on meshes of uneven
   size, produced by the batched ensemble launch (Ensemble) -- several
   launches of `launch_runs` runs each.
2. ``WarmStart.fit``: an MLP (BF16 MFMA kernels) from the gauge series to
   the controls; inputs are the gauge depths minus each gauge's initial
   depth, centred per feature and divided by one common scale; outputs
   log(n / 0.03) per zone and Q / Q_default - 1 per knot, standardised per
   output.
3. ``time_to_misfit``: FP64 4D-Var (Calibration.fourdvar, projected L-BFGS
   on the adjoint gradient) from the default controls ("plain") and from
   the network's prediction for the observed series ("learned"), each run
   until the gauge misfit J falls to the same tolerance; the wall time of
   each (the learned one includes the network's inference) is the figure.

The reference has no learned component; its 4D-Var starts from the
configured controls (assim/swSpE_minimize.cpp:398-511), which is the
"plain" leg here.
"""
import math
import time

import numpy as np

from .ensemble import Autotuner, EnsembleProblem, log_uniform_sizes
from .mlp import MLP

HIDDEN = (512, 512, 512)


def _say(log, msg):
    if log:
        log(msg)


def training_set(n_runs, launch_runs=1024, n_meshes=64, cells=(2.0e4, 6.0e4), window=1200.0,
                 level_quantile=0.9, seed=11, block=None, log=None, device=-1):
    """Gauge series of n_runs runs at prior controls.

    `n_meshes` meshes with cell counts log-uniform in `cells`, each carrying
    launch_runs / n_meshes runs per launch; ceil(n_runs / launch_runs)
    launches with fresh controls.  Returns a dict: gauges [n, n_obs,
    n_gauges] (float32, m), h0 [n, n_gauges] (initial gauge depths), n
    [n, 8], Q [n, 60], and the launch statistics (cell_steps, device_ms,
    wall_s, block)."""
    rep = max(1, launch_runs // n_meshes)
    t0 = time.perf_counter()
    P = EnsembleProblem(sizes=log_uniform_sizes(n_meshes, seed, *cells), window=window,
                        level_quantile=level_quantile, replicate=rep, device=device)
    R = len(P.cals)
    _say(log, f"ensemble: {n_meshes} meshes x {rep} runs = {R} runs per launch, "
              f"{P.n_cells} cells resident ({P.ens.device_bytes() / 1e9:.2f} GB), "
              f"built in {time.perf_counter() - t0:.1f} s")
    if block is None:
        n0, Q0 = P.sample(seed)
        P.set_controls(n0, Q0)
        block = Autotuner().pick(P.ens, P.obs_t)
    h0_run = np.array([c.case.a["h0"][c.gauge_cells] for c in P.cals])
    cells_run = np.array([c.case.n_cell for c in P.cals], dtype=np.int64)
    n_launch = int(math.ceil(n_runs / R))
    G, H, Ns, Qs = [], [], [], []
    cell_steps, dev_ms = 0, 0.0
    for b in range(n_launch):
        n, Q, g, steps, ms = P.run_batch(seed * 1000 + b + 1, block=block)
        G.append(g.astype(np.float32))
        H.append(h0_run)
        Ns.append(n)
        Qs.append(Q)
        cell_steps += int(np.sum(cells_run * steps))
        dev_ms += ms
        _say(log, f"ensemble launch {b + 1}/{n_launch}: {ms / 1e3:.1f} s device, "
                  f"steps {int(steps.min())}..{int(steps.max())}")
    out = dict(gauges=np.concatenate(G)[:n_runs], h0=np.concatenate(H)[:n_runs],
               n=np.concatenate(Ns)[:n_runs], Q=np.concatenate(Qs)[:n_runs],
               cell_steps=cell_steps, device_ms=dev_ms, wall_s=time.perf_counter() - t0,
               block=int(block), runs_per_launch=R, launches=n_launch,
               Q_default=P.cals[0].Q_default, obs_t=P.obs_t)
    P.ens.close()
    return out


class WarmStart:
    """The network and the normalisation of its inputs and outputs."""

    def __init__(self, Q_default, n_ref=0.03, hidden=HIDDEN):
        self.Qd = np.asarray(Q_default, dtype=np.float64)
        self.n_ref = float(n_ref)
        self.hidden = tuple(hidden)
        self.net = None

    # -- encodings -------------------------------------------------------------
    @staticmethod
    def features(gauges, h0):
        """Gauge depths minus the gauge's initial depth, flattened gauge-major."""
        g = np.asarray(gauges, dtype=np.float64) - np.asarray(h0, dtype=np.float64)[..., None, :]
        return g.swapaxes(-1, -2).reshape(g.shape[:-2] + (-1,))

    def targets(self, n, Q):
        return np.concatenate([np.log(np.asarray(n, dtype=np.float64) / self.n_ref),
                               np.asarray(Q, dtype=np.float64) / self.Qd - 1.0], axis=-1)

    def controls(self, y):
        y = np.asarray(y, dtype=np.float64)
        nz = y.shape[-1] - len(self.Qd)
        return self.n_ref * np.exp(y[..., :nz]), self.Qd * (1.0 + y[..., nz:])

    # -- training / inference --------------------------------------------------
    def fit(self, gauges, h0, n, Q, epochs=60, batch=256, lr=1e-3, lr_decay=0.95, val_frac=0.0625,
            seed=0, log=None):
        """Adam on standardised inputs / targets; returns the per-epoch
        (train loss, validation loss) history."""
        X = self.features(gauges, h0)
        Y = self.targets(n, Q)
        rng = np.random.default_rng(seed)
        idx = rng.permutation(len(X))
        n_val = int(round(val_frac * len(X)))
        va, tr = idx[:n_val], idx[n_val:]
        # inputs: centred per feature, ONE scale for all of them -- early
        # samples barely move, and scaling them by their own (tiny) spread
        # would blow the mesh-resolution differences up into the features
        self.x_mu = X[tr].mean(0)
        self.x_sd = float(np.sqrt(np.mean((X[tr] - self.x_mu) ** 2)))
        self.y_mu, self.y_sd = Y[tr].mean(0), np.maximum(Y[tr].std(0), 1e-6)
        xs = ((X - self.x_mu) / self.x_sd).astype(np.float32)
        ys = ((Y - self.y_mu) / self.y_sd).astype(np.float32)
        self.net = MLP([X.shape[1], *self.hidden, Y.shape[1]], max_batch=max(batch, 1024), seed=seed)
        self.net.set_data(xs[tr], ys[tr])
        hist = []
        for ep in range(epochs):
            tl = self.net.train_resident(epochs=1, batch=batch, lr=lr, seed=seed + ep)[0]
            vl = float(np.mean((self.net.forward(xs[va]) - ys[va]) ** 2)) if n_val else float("nan")
            hist.append((tl, vl))
            lr *= lr_decay
            if log and (ep % 10 == 9 or ep == epochs - 1):
                log(f"mlp epoch {ep + 1}/{epochs}: train {tl:.4f} validation {vl:.4f}")
        return hist

    def save(self, path, **extra):
        np.savez(path, params=self.net.params(), dims=np.asarray(self.net.dims), Qd=self.Qd,
                 n_ref=self.n_ref, x_mu=self.x_mu, x_sd=self.x_sd, y_mu=self.y_mu, y_sd=self.y_sd,
                 **extra)

    @classmethod
    def load(cls, path):
        d = np.load(path)
        dims = [int(v) for v in d["dims"]]
        ws = cls(d["Qd"], float(d["n_ref"]), hidden=dims[1:-1])
        ws.x_mu, ws.x_sd, ws.y_mu, ws.y_sd = d["x_mu"], float(d["x_sd"]), d["y_mu"], d["y_sd"]
        ws.net = MLP(dims, max_batch=1024)
        ws.net.set_params(d["params"])
        return ws, {k: d[k] for k in d.files}

    def predict(self, gauges, h0):
        """Controls (n [.., 8], Q [.., 60]) for observed gauge series."""
        x = ((self.features(gauges, h0) - self.x_mu) / self.x_sd).astype(np.float32)
        one = x.ndim == 1
        y = self.net.forward(np.atleast_2d(x)).astype(np.float64) * self.y_sd + self.y_mu
        n, Q = self.controls(y)
        return (n[0], Q[0]) if one else (n, Q)


def time_to_misfit(cal, solver, obs, starts, J_tol, max_iter=200, time_limit=600.0,
                   max_steps=30000, n_tiles=1, adj_cut=0.0, log=None):
    """4D-Var from each start in `starts` ({name: (n0, Q0, setup_s)}) until
    J <= J_tol; per start: seconds to the tolerance (None if not reached,
    setup_s included), gradient and cost evaluations, the J trace, the final
    controls' distance to the truth."""
    res = {}
    for name, (n0, Q0, setup_s) in starts.items():
        n, Q, trace = cal.fourdvar(solver, obs, n0, Q0, J_tol=J_tol, max_iter=max_iter,
                                   max_steps=max_steps, time_limit=time_limit, n_tiles=n_tiles,
                                   max_ls=12, adj_cut=adj_cut, log=log)
        hit = next((t for t, J, _ in trace if J <= J_tol), None)
        res[name] = dict(
            seconds=None if hit is None else hit + setup_s, setup_s=setup_s,
            J0=trace[0][1], J_final=min(J for _, J, _ in trace), wall_s=trace[-1][0] + setup_s,
            grads=sum(k == "grad" for *_, k in trace), costs=sum(k == "ls" for *_, k in trace),
            n_err=float(np.max(np.abs(n / cal.n_true - 1.0))),
            Q_err=float(np.sqrt(np.mean((Q - cal.Q_true) ** 2)) / np.mean(cal.Q_true)),
            trace=[(round(t, 3), float(J), k) for t, J, k in trace])
        _say(log, f"4D-Var {name}: J {res[name]['J0']:.3e} -> {res[name]['J_final']:.3e} in "
                  f"{res[name]['wall_s']:.1f} s ({res[name]['grads']} gradients); to J <= {J_tol:.2e}: "
                  f"{res[name]['seconds']}")
    return res
