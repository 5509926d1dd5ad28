"""Batched ensemble of differentiable-forward runs over uneven meshes
(include/sw2dg.h, sw2dg_ens_*), its autotuner, and the ensemble workload of
the calibration problem.

``Ensemble(cases)`` holds many meshes on the GPU; ``set_run`` gives a run its
controls, hydrograph, initial state and gauges; ``run`` advances every run to
t_max in one persistent launch and returns the gauge series of every run.
Run r reproduces Assim.record on its own mesh bit for bit.

``Autotuner`` picks the workgroup size of the launch (the block of edges or
cells a workgroup processes per pass) by timing a few steps of each candidate
on the actual ensemble, and remembers the choice per ensemble shape.
"""
import ctypes as C
import json
import os

import numpy as np

from . import _native as N

_DECL = {
    "sw2dg_ens_create": ([C.c_int32, C.c_void_p, C.POINTER(N.Params), C.c_int32,
                          C.POINTER(C.c_void_p)], C.c_int),
    "sw2dg_ens_set_run": ([C.c_void_p, C.c_int32, N._d, N._d, N._d, N._d, C.c_int32, C.c_int32,
                           N._i32, N._d, N._d, N._d, N._d, N._d, C.c_int32, N._i64], C.c_int),
    "sw2dg_ens_run": ([C.c_void_p, C.c_int32, N._d, C.c_int32, C.c_int64, N._d, N._i64,
                       C.POINTER(C.c_double)], C.c_int),
    "sw2dg_ens_get_state": ([C.c_void_p, C.c_int32, N._d, N._d, N._d], C.c_int),
    "sw2dg_ens_set_cluster_limit": ([C.c_void_p, C.c_int32], C.c_int),
    "sw2dg_ens_last_launch": ([C.c_void_p, N._i32, N._i32], C.c_int),
    "sw2dg_ens_device_bytes": ([C.c_void_p], C.c_int64),
    "sw2dg_ens_last_error": ([C.c_void_p], C.c_char_p),
    "sw2dg_ens_destroy": ([C.c_void_p], None),
}
ENS_SYMBOLS = tuple(_DECL)
BLOCKS = (256, 512, 1024)


def _lib():
    lib = N.gpu_lib()
    if not getattr(lib, "_ens_ready", False):
        for k, (a, r) in _DECL.items():
            f = getattr(lib, k)
            f.argtypes = a
            f.restype = r
        lib._ens_ready = True
    return lib


class Ensemble:
    """Meshes of many runs resident on one GPU (one sw2dg_ens)."""

    def __init__(self, cases, params=None, device=-1):
        self.lib = _lib()
        self.cases = list(cases)
        self.n_runs = len(self.cases)
        self._keep = []
        meshes = (N.Mesh * self.n_runs)()
        for r, case in enumerate(self.cases):
            a = case.a
            m = meshes[r]
            m.n_cell, m.n_bnd, m.n_int = case.n_cell, case.n_bnd, case.n_int
            for f in ("cell_area", "cell_x", "cell_y", "cell_zb", "bnd_len", "bnd_nx", "bnd_ny",
                      "int_len", "int_nx", "int_ny", "int_xm", "int_ym"):
                setattr(m, f, N.dptr(self._hold(a[f], np.float64)))
            m.bnd_cell = N.i64ptr(self._hold(a["bnd_cell"], np.int64))
            m.int_L = N.i64ptr(self._hold(a["int_L"], np.int64))
            m.int_R = N.i64ptr(self._hold(a["int_R"], np.int64))
            m.bnd_series = N.i32ptr(self._hold(a["bnd_series"], np.int32))
        self.params = params if params is not None else self.cases[0].params
        h = C.c_void_p()
        rc = self.lib.sw2dg_ens_create(self.n_runs, C.cast(meshes, C.c_void_p), C.byref(self.params),
                                       device, C.byref(h))
        if rc != 0:
            raise N.SW2DError(rc, "sw2dg_ens_create failed")
        self.h = h
        self._meshes = meshes
        self.n_gauges = np.zeros(self.n_runs, dtype=np.int64)

    def _hold(self, x, dt):
        x = np.ascontiguousarray(x, dtype=dt)
        self._keep.append(x)
        return x

    def _check(self, rc):
        if rc != 0:
            raise N.SW2DError(rc, self.lib.sw2dg_ens_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.sw2dg_ens_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_run(self, r, phi, K1, K2, types, t, v, h, qx, qy, gauge_cells, fric_alpha_deg=None):
        nc = self.cases[r].n_cell
        f = lambda x: np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (nc,)))
        v = np.atleast_2d(np.asarray(v, dtype=np.float64))
        ty = np.ascontiguousarray(types, dtype=np.int32)
        t = np.ascontiguousarray(t, dtype=np.float64)
        vv = np.ascontiguousarray(v.reshape(-1))
        gc = np.ascontiguousarray(gauge_cells, dtype=np.int64)
        arrs = [f(phi), f(K1), f(K2), f(h), f(qx), f(qy)]
        fa = None if fric_alpha_deg is None else f(fric_alpha_deg)
        self._check(self.lib.sw2dg_ens_set_run(
            self.h, r, N.dptr(arrs[0]), N.dptr(arrs[1]), N.dptr(arrs[2]),
            None if fa is None else N.dptr(fa), v.shape[0], len(t), N.i32ptr(ty), N.dptr(t),
            N.dptr(vv), N.dptr(arrs[3]), N.dptr(arrs[4]), N.dptr(arrs[5]), len(gc), N.i64ptr(gc)))
        self.n_gauges[r] = len(gc)

    def run(self, obs_t, block=256, max_steps=10 ** 9):
        """Every run to t_max: (gauges [n_runs, n_obs, max n_gauges], steps [n_runs], device ms)."""
        obs_t = np.ascontiguousarray(obs_t, dtype=np.float64)
        ng = max(int(self.n_gauges.max()), 1)
        g = np.zeros(self.n_runs * max(len(obs_t), 1) * ng)
        steps = np.zeros(self.n_runs, dtype=np.int64)
        ms = C.c_double()
        self._check(self.lib.sw2dg_ens_run(self.h, len(obs_t), N.dptr(obs_t), int(block),
                                           int(max_steps), N.dptr(g), N.i64ptr(steps),
                                           C.byref(ms)))
        return g.reshape(self.n_runs, max(len(obs_t), 1), ng)[:, : len(obs_t)], steps, ms.value

    def state(self, r):
        nc = self.cases[r].n_cell
        h, qx, qy = np.empty(nc), np.empty(nc), np.empty(nc)
        self._check(self.lib.sw2dg_ens_get_state(self.h, r, N.dptr(h), N.dptr(qx), N.dptr(qy)))
        return h, qx, qy

    def device_bytes(self):
        return self.lib.sw2dg_ens_device_bytes(self.h)

    def set_cluster_limit(self, max_wg_per_run):
        """Workgroups a run may get when all runs fit the GPU at once (1: one each)."""
        self._check(self.lib.sw2dg_ens_set_cluster_limit(self.h, int(max_wg_per_run)))

    def last_launch(self):
        """(workgroups, largest cluster) of the last run()."""
        w, c = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
        self._check(self.lib.sw2dg_ens_last_launch(self.h, N.i32ptr(w), N.i32ptr(c)))
        return int(w[0]), int(c[0])


class Autotuner:
    """Picks the workgroup size of the ensemble launch: every candidate runs
    `probe_steps` steps of the actual ensemble (the largest runs set the
    pace), the fastest per step wins; choices are kept per ensemble shape
    (number of runs and total cells, rounded) in memory and, if `path` is
    given, in a JSON file."""

    def __init__(self, path=None, candidates=BLOCKS, probe_steps=40):
        self.path = path
        self.candidates = tuple(candidates)
        self.probe_steps = int(probe_steps)
        self.table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)
        self.last = {}

    @staticmethod
    def key(ens):
        cells = sum(c.n_cell for c in ens.cases)
        big = max(c.n_cell for c in ens.cases)
        return f"runs{ens.n_runs}_cells{int(round(cells, -4))}_max{int(round(big, -3))}"

    def pick(self, ens, obs_t):
        k = self.key(ens)
        if k in self.table:
            return int(self.table[k]["block"])
        timing = {}
        for b in self.candidates:
            _, steps, ms = ens.run(obs_t, block=b, max_steps=self.probe_steps)
            timing[b] = ms / max(int(steps.max()), 1)
        best = min(timing, key=timing.get)
        self.table[k] = {"block": int(best), "ms_per_step": {str(b): v for b, v in timing.items()}}
        self.last = timing
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1)
        return int(best)


def log_uniform_sizes(n_runs, seed, lo=2.0e4, hi=2.5e5):
    """Lattice sizes px (2 px^2 triangles) of n_runs meshes with cell counts
    log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    cells = np.exp(rng.uniform(np.log(lo), np.log(hi), n_runs))
    return np.maximum(4, np.round(np.sqrt(cells / 2.0)).astype(int))


class EnsembleProblem:
    """The calibration problem on n_runs meshes of the same domain (sizes
    log-uniform), each a Calibration; batches of seeded control sets from the
    problem's prior run in one launch and return the gauge series."""

    def __init__(self, n_runs=256, seed=7, window=1200.0, level_quantile=0.9, sizes=None,
                 device=-1, replicate=1, workers=None):
        from concurrent.futures import ThreadPoolExecutor

        from .calibration import Calibration
        px = log_uniform_sizes(n_runs, seed) if sizes is None else np.asarray(sizes)
        # the meshes are independent and deterministic: built on `workers`
        # host threads (the Delaunay and zoning work releases the GIL)
        workers = workers or min(8, os.cpu_count() or 1)
        build = lambda p: Calibration(px=int(p), window=window, level_quantile=level_quantile)
        with ThreadPoolExecutor(max_workers=workers) as ex:
            meshes = list(ex.map(build, px))
        # `replicate` runs per mesh (distinct controls, one device copy each)
        self.cals = [c for c in meshes for _ in range(int(replicate))]
        self.ens = Ensemble([c.case for c in self.cals], device=device)
        c0 = self.cals[0]
        self.obs_t = c0.obs_t
        self.n_zones, self.n_knots = c0.n_zones, len(c0.t_knots)

    @property
    def n_cells(self):
        return int(sum(c.case.n_cell for c in self.cals))

    def set_controls(self, n_zone, Q):
        """n_zone [n_runs, 8], Q [n_runs, 60]."""
        types = np.array([N.BC_Q, N.BC_Q], dtype=np.int32)
        for r, c in enumerate(self.cals):
            a = c.case.a
            K = c.K(n_zone[r])
            self.ens.set_run(r, a["phi"], K, K, types, c.t_knots, c.bc_values(Q[r]), a["h0"],
                             a["qx0"], a["qy0"], c.gauge_cells)

    def sample(self, seed):
        """Controls of every run drawn from the prior (Calibration.sample_controls)."""
        n = np.empty((len(self.cals), self.n_zones))
        Q = np.empty((len(self.cals), self.n_knots))
        for r, c in enumerate(self.cals):
            n[r], Q[r] = c.sample_controls(int(seed) * 100003 + r)
        return n, Q

    def run_batch(self, seed, block=256):
        """One launch at seeded controls: (n, Q, gauges [n_runs, n_obs, n_gauges], steps, ms)."""
        n, Q = self.sample(seed)
        self.set_controls(n, Q)
        g, steps, ms = self.ens.run(self.obs_t, block=block)
        return n, Q, g, steps, ms
