"""4D-Var layer: host mirror of the reference's swSpE assimilation entry points.

``Assim`` wraps the C-ABI of include/sw2dg.h (sw2dg_assim_*) on an SpSolver
context:
  record()   ~ swSpE::build_checkpoint_records  (assim/swSpE_checkpoint.cpp:97-455)
  cost()     ~ swSpE::replay_trajectory_cost    (:1311-1433)
  gradient() ~ swSpE::get_spatial_gradient_checkpointed (:722-1281)
All of them run the _E forward and the hand-written adjoint on the GPU.

``Controls`` maps the 68 controls of the calibration workload (Manning
roughness of 8 zones and a 60-knot inflow hydrograph) to the per-cell /
per-step inputs of the kernels, and the per-cell / per-step gradients back.
The friction of the _E forward is the semi-implicit Strickler law
(swSpE_friction.cpp:19-76), so a Manning n enters as K = 1/n.
"""
import ctypes as C

import numpy as np

from . import _native as N


class Obs(C.Structure):
    """sw2dg_obs (include/sw2dg.h)."""

    _fields_ = [
        ("n_obs", C.c_int32), ("t", N._d), ("n_gauges", C.c_int32), ("gauge_cell", N._i64),
        ("gauge_w", N._d), ("gauge_h", N._d), ("frame_h", N._d), ("frame_qx", N._d),
        ("frame_qy", N._d), ("w_h", N._d), ("w_q", N._d), ("normalize", C.c_int32),
    ]


_DECL = {
    "sw2dg_assim_setup": ([C.c_void_p, C.POINTER(Obs), C.c_int32, C.c_int32, C.c_double, C.c_int64], C.c_int),
    "sw2dg_assim_record": ([C.c_void_p, N._d, N._d, N._d, N._i64], C.c_int),
    "sw2dg_assim_get_record": ([C.c_void_p, N._d, N._d, N._d, N._i32], C.c_int),
    "sw2dg_assim_set_bc_seq": ([C.c_void_p, N._d], C.c_int),
    "sw2dg_assim_cost": ([C.c_void_p, N._d, N._d, N._d, N._d], C.c_int),
    "sw2dg_assim_gradient": ([C.c_void_p, N._d, N._d, N._d, N._d, N._d, N._d, N._d, N._d, N._i64],
                             C.c_int),
    "sw2dg_assim_final_state": ([C.c_void_p, N._d, N._d, N._d], C.c_int),
    "sw2dg_assim_gauge_series": ([C.c_void_p, N._d], C.c_int),
    "sw2dg_assim_keep_obs_states": ([C.c_void_p, C.c_int32], C.c_int),
    "sw2dg_assim_adjoint_peak": ([C.c_void_p, N._d], C.c_int),
    "sw2dg_assim_obs_states": ([C.c_void_p, N._d, N._d, N._d, N._i32], C.c_int),
    "sw2dg_assim_set_adjoint_cut": ([C.c_void_p, C.c_double], C.c_int),
    "sw2dg_assim_set_explosion_guard": ([C.c_void_p, C.c_double], C.c_int),
    "sw2dg_assim_set_resident": ([C.c_void_p, C.c_int32], C.c_int),
    "sw2dg_assim_resident": ([C.c_void_p, N._i32], C.c_int),
    "sw2dg_assim_set_phi_transform": ([C.c_void_p, C.c_int32, C.c_double, N._i64, C.c_int64,
                                       C.c_int32], C.c_int),
    "sw2dg_assim_truncation": ([C.c_void_p, N._i64, N._i64], C.c_int),
    "sw2dg_assim_set_first_segment": ([C.c_void_p, C.c_int64], C.c_int),
    "sw2dg_assim_profile": ([C.c_void_p, N._d, N._d, N._d, C.c_int32, N._d, N._i64,
                             C.POINTER(C.c_char_p), C.POINTER(C.c_int32)], C.c_int),
}
ASSIM_SYMBOLS = tuple(_DECL) + ("sw2dg_phi_transform_eval",)
PHI_MODES = {None: 0, "none": 0, "clamp": 1, "tanh": 2, "normalize": 3}


def phi_transform_eval(mode, m, lam=1.0, lo=0.0, hi=1.0):
    """phi_correct and dphi/dm of the library (host code, no device): (phi, dphi)."""
    lib = N.gpu_lib()
    f = lib.sw2dg_phi_transform_eval
    f.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int64, N._d, N._d, N._d]
    f.restype = C.c_int
    m = np.ascontiguousarray(m, dtype=np.float64)
    phi, dphi = np.empty_like(m), np.empty_like(m)
    rc = f(PHI_MODES[mode], float(lam), float(lo), float(hi), len(m), N.dptr(m), N.dptr(phi),
           N.dptr(dphi))
    if rc != 0:
        raise N.SW2DError(rc, "sw2dg_phi_transform_eval failed")
    return phi, dphi


def _declare(lib):
    for k, (a, r) in _DECL.items():
        f = getattr(lib, k)
        f.argtypes = a
        f.restype = r


class Assim:
    """Frozen-trajectory cost and adjoint on one SpSolver context."""

    def __init__(self, solver, obs_t, gauge_cells=None, gauge_h=None, gauge_w=None, frames=None,
                 frame_w=None, normalize=False, stride=0, n_tiles=1, t_window=-1e300,
                 max_steps=100000, adj_cut=0.0, explosion_ratio=1e8, phi_transform=None,
                 free_cells=None, mask_wall=False, resident=True):
        self.s = solver
        self.lib = solver.lib
        _declare(self.lib)
        self._keep = []
        self.max_steps = int(max_steps)
        o = Obs()
        obs_t = self._hold(np.atleast_1d(np.asarray(obs_t, dtype=np.float64)))
        o.n_obs = len(obs_t)
        o.t = N.dptr(obs_t)
        if gauge_cells is not None:
            gc = self._hold(np.asarray(gauge_cells, dtype=np.int64))
            o.n_gauges = len(gc)
            o.gauge_cell = N.i64ptr(gc)
            o.gauge_h = N.dptr(self._hold(np.asarray(gauge_h, dtype=np.float64).reshape(-1)))
            if gauge_w is not None:
                o.gauge_w = N.dptr(self._hold(np.asarray(gauge_w, dtype=np.float64)))
        if frames is not None:
            fh, fqx, fqy = (self._hold(np.asarray(f, dtype=np.float64).reshape(-1)) for f in frames)
            o.frame_h, o.frame_qx, o.frame_qy = N.dptr(fh), N.dptr(fqx), N.dptr(fqy)
            if frame_w is not None:
                o.w_h = N.dptr(self._hold(np.asarray(frame_w[0], dtype=np.float64)))
                o.w_q = N.dptr(self._hold(np.asarray(frame_w[1], dtype=np.float64)))
        o.normalize = 1 if normalize else 0
        self.obs = o
        self.n_obs = o.n_obs
        self.s._check(self.lib.sw2dg_assim_setup(solver.ctx, C.byref(o), int(stride), int(n_tiles),
                                                 float(t_window), self.max_steps))
        # the cut is a property of the solver context: set it on every Assim
        self.set_adjoint_cut(adj_cut)
        self.s._check(self.lib.sw2dg_assim_set_explosion_guard(solver.ctx, float(explosion_ratio)))
        # resident trajectory (Godunov gradients read the dry run's steps in
        # place of replaying them, where device memory allows)
        self.s._check(self.lib.sw2dg_assim_set_resident(solver.ctx, 1 if resident else 0))
        self.stride = int(stride) if stride else int(np.ceil(np.sqrt(self.max_steps)))
        # phi_transform: None or (mode, lambda), mode "clamp" | "tanh" | "normalize";
        # the phi passed to record / cost / gradient is then the control m
        mode, lam = (None, 1.0) if phi_transform is None else (tuple(phi_transform) + (1.0,))[:2]
        code = PHI_MODES[mode]
        fc = np.ascontiguousarray([] if free_cells is None else free_cells, dtype=np.int64)
        self._free = fc
        self.s._check(self.lib.sw2dg_assim_set_phi_transform(
            solver.ctx, code, float(lam), N.i64ptr(fc) if len(fc) else None, len(fc),
            1 if mask_wall else 0))
        self.n_steps = 0
        self._tmp = []

    def set_adjoint_cut(self, cut):
        """Thin-film cut of the adjoint (m; 0 = the exact adjoint) for the
        following gradients (sw2dg_assim_set_adjoint_cut)."""
        self.adj_cut = float(cut)
        self.s._check(self.lib.sw2dg_assim_set_adjoint_cut(self.s.ctx, self.adj_cut))

    def _hold(self, a):
        a = np.ascontiguousarray(a)
        self._keep.append(a)
        return a

    def _ctl(self, x):
        """Per-call control array (the library copies it before returning, so
        it lives until the next call only)."""
        if x is None:
            return None
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64),
                                                 (self.s.case.n_cell,)))
        self._tmp.append(a)
        return N.dptr(a)

    def record(self, phi=None, K1=None, K2=None):
        self._tmp = []
        n = C.c_int64()
        self.s._check(self.lib.sw2dg_assim_record(self.s.ctx, self._ctl(phi), self._ctl(K1),
                                                  self._ctl(K2), C.byref(n)))
        self.n_steps = n.value
        return self.n_steps

    def get_record(self):
        n = self.n_steps
        ns = max(self.s.case.n_series, 1)
        t = np.empty(n + 1)
        dt = np.empty(max(n, 1))
        bc = np.empty(max(n, 1) * ns)
        trig = np.empty(max(n, 1), dtype=np.int32)
        self.s._check(self.lib.sw2dg_assim_get_record(self.s.ctx, N.dptr(t), N.dptr(dt),
                                                      N.dptr(bc), N.i32ptr(trig)))
        return {"t": t, "dt": dt[:n], "bc": bc[: n * ns].reshape(n, ns), "trig": trig[:n]}

    def set_bc_seq(self, bc):
        bc = np.ascontiguousarray(np.asarray(bc, dtype=np.float64).reshape(-1))
        self.s._check(self.lib.sw2dg_assim_set_bc_seq(self.s.ctx, N.dptr(bc)))

    def cost(self, phi=None, K1=None, K2=None):
        self._tmp = []
        J = C.c_double()
        self.s._check(self.lib.sw2dg_assim_cost(self.s.ctx, self._ctl(phi), self._ctl(K1),
                                                self._ctl(K2), C.byref(J)))
        return J.value

    def gradient(self, phi=None, K1=None, K2=None):
        """Dry run at the controls, then the adjoint.  Returns J and dJ/d(phi, K1, K2, bc)."""
        self._tmp = []
        nc = self.s.case.n_cell
        ns = max(self.s.case.n_series, 1)
        J = C.c_double()
        n = C.c_int64()
        gphi, gK1, gK2 = np.empty(nc), np.empty(nc), np.empty(nc)
        gbc = np.empty(self.max_steps * ns)
        self.s._check(self.lib.sw2dg_assim_gradient(self.s.ctx, self._ctl(phi), self._ctl(K1),
                                                    self._ctl(K2), C.byref(J), N.dptr(gphi),
                                                    N.dptr(gK1), N.dptr(gK2), N.dptr(gbc),
                                                    C.byref(n)))
        self.n_steps = n.value
        return {"J": J.value, "phi": gphi, "K1": gK1, "K2": gK2,
                "bc": gbc[: n.value * ns].reshape(n.value, ns), "n_steps": n.value}

    def resident_segments(self):
        """Segments the last gradient read from resident blocks (0: every
        segment was replayed from its checkpoint)."""
        v = np.zeros(1, dtype=np.int32)
        self.s._check(self.lib.sw2dg_assim_resident(self.s.ctx, N.i32ptr(v)))
        return int(v[0])

    def truncation(self):
        """(first kept step, first kept segment) of the last gradient: (0, 0)
        unless the explosion guard stopped the reverse sweep."""
        st, sg = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
        self.s._check(self.lib.sw2dg_assim_truncation(self.s.ctx, N.i64ptr(st), N.i64ptr(sg)))
        return int(st[0]), int(sg[0])

    def set_first_segment(self, segment):
        """Freeze the truncation window at `segment` (the reference's
        forced_first_seg_): later gradients and costs start there, with the
        explosion guard's dynamic truncation off; None or < 0 restores the
        dynamic truncation."""
        seg = -1 if segment is None else int(segment)
        self.s._check(self.lib.sw2dg_assim_set_first_segment(self.s.ctx, seg))

    def profile(self, phi=None, K1=None, K2=None, k_max=32):
        """One gradient with per-launch HIP events: {kernel: (avg_ms, launches)}."""
        self._tmp = []
        avg = np.zeros(k_max)
        cnt = np.zeros(k_max, dtype=np.int64)
        names = (C.c_char_p * k_max)()
        nk = C.c_int32()
        self.s._check(self.lib.sw2dg_assim_profile(self.s.ctx, self._ctl(phi), self._ctl(K1),
                                                   self._ctl(K2), k_max, N.dptr(avg),
                                                   N.i64ptr(cnt), names, C.byref(nk)))
        return {names[k].decode(): (float(avg[k]), int(cnt[k])) for k in range(min(nk.value, k_max))}

    def gauge_series(self):
        """[n_obs, n_gauges] depths at the gauge cells sampled by the last dry run."""
        ng = self.obs.n_gauges
        out = np.zeros(max(self.n_obs * ng, 1))
        self.s._check(self.lib.sw2dg_assim_gauge_series(self.s.ctx, N.dptr(out)))
        return out[: self.n_obs * ng].reshape(self.n_obs, ng)

    def adjoint_peak(self):
        """Largest ||adjoint of h|| at a segment start of the last gradient."""
        v = C.c_double()
        self.s._check(self.lib.sw2dg_assim_adjoint_peak(self.s.ctx, C.byref(v)))
        return v.value

    def keep_obs_states(self, on=True):
        """Keep the full state of every observation step of later dry runs."""
        self.s._check(self.lib.sw2dg_assim_keep_obs_states(self.s.ctx, 1 if on else 0))

    def obs_states(self):
        """(h, qx, qy) [n_obs, n_cell] at the observation steps of the last dry
        run, and reached[n_obs] (False where the run did not get there)."""
        nc = self.s.case.n_cell
        h, qx, qy = (np.empty(self.n_obs * nc) for _ in range(3))
        reached = np.zeros(self.n_obs, dtype=np.int32)
        self.s._check(self.lib.sw2dg_assim_obs_states(self.s.ctx, N.dptr(h), N.dptr(qx), N.dptr(qy),
                                                      N.i32ptr(reached)))
        sh = (self.n_obs, nc)
        return h.reshape(sh), qx.reshape(sh), qy.reshape(sh), reached.astype(bool)

    def final_state(self):
        nc = self.s.case.n_cell
        h, qx, qy = np.empty(nc), np.empty(nc), np.empty(nc)
        self.s._check(self.lib.sw2dg_assim_final_state(self.s.ctx, N.dptr(h), N.dptr(qx),
                                                       N.dptr(qy)))
        return h, qx, qy


def interp_weights(tk, t):
    """Weights W[n, k] of the _E time interpolation (assim/swSpE.h:138-153):
    value(t_n) = sum_k W[n, k] v_k.  Used to map boundary-value gradients per
    step onto hydrograph knots."""
    tk = np.asarray(tk, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    W = np.zeros((len(t), len(tk)))
    nk = len(tk)
    for n, tn in enumerate(t):
        if nk == 1 or tn <= tk[0] - 1e-12:
            W[n, 0] = 1.0
            continue
        if tn >= tk[-1]:
            W[n, -1] = 1.0
            continue
        it = 0
        while it < nk - 1 and tn > tk[it + 1]:
            it += 1
        den = tk[it + 1] - tk[it]
        if den < 1e-12:
            W[n, it] = 1.0
            continue
        a = (tn - tk[it]) / den
        W[n, it] += 1.0 - a
        W[n, it + 1] += a
    return W


class Controls:
    """Manning zones + inflow hydrograph <-> per-cell K and per-step BC values.

    zone[c] in [0, n_zones); hydrograph knots (t_k, Q_k) in m3/s on the
    boundary series `series` whose edges have total length `width` (unit
    discharge = Q / width)."""

    def __init__(self, zone, n_zones, t_knots, series, width):
        self.zone = np.asarray(zone, dtype=np.int64)
        self.n_zones = int(n_zones)
        self.t_knots = np.asarray(t_knots, dtype=np.float64)
        self.series = int(series)
        self.width = float(width)

    def K(self, n_zone):
        return 1.0 / np.asarray(n_zone, dtype=np.float64)[self.zone]

    def grad_zones(self, n_zone, dK1, dK2):
        """dJ/dn_z from dJ/dK per cell (K1 = K2 = 1/n)."""
        n_zone = np.asarray(n_zone, dtype=np.float64)
        g = np.zeros(self.n_zones)
        np.add.at(g, self.zone, dK1 + dK2)
        return g * (-1.0 / n_zone ** 2)

    def grad_knots(self, t_steps, dbc):
        """dJ/dQ_k from dJ/d(boundary value) per step (frozen step times)."""
        W = interp_weights(self.t_knots, t_steps)
        return (W.T @ dbc[:, self.series]) / self.width
