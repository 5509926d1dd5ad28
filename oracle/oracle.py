"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the checker for the HIP path.  Only
tests, the smoke check and the benchmark's CPU-baseline leg use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
# the same restatement built with the kernels' exp / log / cbrt (orc_math.h)
LIB_DM = os.path.join(HERE, "_build", "liboracle_dm.so")
# the glibc restatement's OpenMP build (orc_omp.h): the all-core CPU baseline
LIB_OMP = os.path.join(HERE, "_build", "liboracle_omp.so")
MATHS = ("glibc", "device", "omp")
REF_SOLVER = os.path.join(HERE, "_ref", "sw2dSolver")

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)


class OrcCase(C.Structure):
    _fields_ = [
        ("n_cell", C.c_int64), ("n_bnd", C.c_int64), ("n_int", C.c_int64),
        ("cell_area", _d), ("cell_x", _d), ("cell_y", _d), ("cell_zb", _d),
        ("phi", _d), ("beta", _d),
        ("fric_c1", _d), ("fric_c2", _d), ("fric_alpha_deg", _d),
        ("hl_hsing", _d), ("hl_c1", _d), ("hl_c2", _d), ("hl_alpha_deg", _d),
        ("bnd_cell", _i64), ("bnd_len", _d), ("bnd_nx", _d), ("bnd_ny", _d), ("bnd_series", _i32),
        ("int_L", _i64), ("int_R", _i64), ("int_len", _d), ("int_nx", _d), ("int_ny", _d),
        ("int_xm", _d), ("int_ym", _d),
        ("n_series", C.c_int32), ("n_times", C.c_int32),
        ("series_type", _i32), ("series_t", _d), ("series_v", _d),
        ("scheme", C.c_int32), ("topo", C.c_int32), ("friction", C.c_int32),
        ("head_loss", C.c_int32), ("neg_depth_detection", C.c_int32),
        ("h_min", C.c_double), ("phi_min", C.c_double), ("cfl_max", C.c_double),
        ("dt_max", C.c_double), ("dt_map", C.c_double), ("t0", C.c_double), ("t_max", C.c_double),
        ("v_max", C.c_double), ("fr_max", C.c_double),
        ("n_wind_station", C.c_int32), ("n_wind_times", C.c_int32),
        ("wind_code", _i32), ("wind_t", _d), ("wind_x", _d), ("wind_y", _d),
        ("runoff_model", C.c_int32), ("rain_water_only", C.c_int32),
        ("n_rain_station", C.c_int32), ("n_rain_times", C.c_int32),
        ("rain_code", _i32), ("rain_t", _d), ("rain_v", _d), ("runoff_coeff", _d),
        ("infil_model", C.c_int32), ("infil_rate", _d), ("momentum_diffusion", C.c_int32),
    ]


_libs = {}


def lib(math="glibc"):
    """The oracle library: math "glibc" (the reference's exp / log / cbrt),
    "device" (the kernels' own, every call audited against glibc) or "omp"
    (glibc's, the OpenMP build: OMP_NUM_THREADS threads)."""
    if math not in MATHS:
        raise ValueError(f"math must be one of {MATHS}")
    if math not in _libs:
        path = {"glibc": LIB, "device": LIB_DM, "omp": LIB_OMP}[math]
        if not os.path.exists(path):
            subprocess.check_call(["make", "-C", HERE], stdout=subprocess.DEVNULL)
        L = C.CDLL(path)
        L.orc_create.argtypes = [C.POINTER(OrcCase), _d, _d, _d]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p]
        L.orc_step.restype = C.c_int
        L.orc_run.argtypes = [C.c_void_p, C.c_double, C.c_int64, _i64]
        L.orc_run.restype = C.c_int
        L.orc_get_state.argtypes = [C.c_void_p, _d, _d, _d]
        L.orc_set_state.argtypes = [C.c_void_p, _d, _d, _d]
        L.orc_time.argtypes = [C.c_void_p]
        L.orc_time.restype = C.c_double
        L.orc_last_dt.argtypes = [C.c_void_p]
        L.orc_last_dt.restype = C.c_double
        L.orc_get_bnd_f1.argtypes = [C.c_void_p, _d]
        L.orc_hllc.argtypes = [C.c_int, C.c_double, C.c_double] + [C.c_double] * 16 + [_d, C.POINTER(C.c_int)]
        L.orc_bnd_flux.argtypes = [C.c_double] * 8 + [C.c_int, C.c_double, _d, C.POINTER(C.c_int)]
        L.orc_interp_series.argtypes = [_d, _d, C.c_int64, C.c_double]
        L.orc_interp_series.restype = C.c_double
        L.orc_threads.restype = C.c_int
        L.orc_friction_factor.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_friction_factor.restype = C.c_double
        L.orc_math_audit_reset.argtypes = []
        L.orc_math_audit_reset.restype = None
        L.orc_math_audit.argtypes = [_i64, _i32, _d, _d, _d, _i64]
        L.orc_math_audit.restype = C.c_int
        L.orc_math_eval.argtypes = [C.c_int32, C.c_int64, _d, _d]
        L.orc_math_eval.restype = None
        _libs[math] = L
    return _libs[math]


def math_eval(fn, x, math="glibc"):
    """The oracle's exp / log / cbrt (fn "exp" | "log" | "cbrt") of an array."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib(math).orc_math_eval({"exp": 0, "log": 1, "cbrt": 2}[fn], len(x), _dp(x), _dp(y))
    return y


def math_audit_reset():
    lib("device").orc_math_audit_reset()


def math_audit():
    """First transcendental call of the device-math oracle whose result
    differs from glibc's since the last reset: dict(step, fn, x, glibc, device,
    count), step -1 / count 0 when none did."""
    st, fn, cnt = C.c_int64(), C.c_int32(), C.c_int64()
    x, g, d = C.c_double(), C.c_double(), C.c_double()
    lib("device").orc_math_audit(C.byref(st), C.byref(fn), C.byref(x), C.byref(g), C.byref(d),
                                 C.byref(cnt))
    return dict(step=st.value, fn=("exp", "log", "cbrt")[fn.value] if fn.value >= 0 else None,
                x=x.value, glibc=g.value, device=d.value, count=cnt.value)


def _dp(a):
    return None if a is None or len(a) == 0 else a.ctypes.data_as(_d)


def build_case(case, params=None, keep=None):
    """OrcCase view of a loaded case (arrays kept alive in `keep`)."""
    a = case.a
    p = params if params is not None else case.params
    keep = keep if keep is not None else []

    def k(arr, dt=np.float64):
        arr = np.ascontiguousarray(arr, dtype=dt)
        keep.append(arr)
        return arr

    c = OrcCase()
    c.n_cell, c.n_bnd, c.n_int = case.n_cell, case.n_bnd, case.n_int
    for f in ("cell_area", "cell_x", "cell_y", "cell_zb", "phi", "beta", "bnd_len", "bnd_nx",
              "bnd_ny", "int_len", "int_nx", "int_ny", "int_xm", "int_ym"):
        setattr(c, f, _dp(k(a[f])))
    for f, src in (("fric_c1", "fc1"), ("fric_c2", "fc2"), ("fric_alpha_deg", "falpha"),
                   ("hl_hsing", "hsing"), ("hl_c1", "hk1"), ("hl_c2", "hk2"),
                   ("hl_alpha_deg", "halpha")):
        setattr(c, f, _dp(k(a[src])))
    c.bnd_cell = k(a["bnd_cell"], np.int64).ctypes.data_as(_i64)
    c.bnd_series = k(a["bnd_series"], np.int32).ctypes.data_as(_i32)
    c.int_L = k(a["int_L"], np.int64).ctypes.data_as(_i64)
    c.int_R = k(a["int_R"], np.int64).ctypes.data_as(_i64)
    c.n_series, c.n_times = case.n_series, case.n_times
    c.series_type = k(a["series_type"], np.int32).ctypes.data_as(_i32)
    c.series_t = _dp(k(a["series_t"]))
    c.series_v = _dp(k(a["series_v"]))
    for f in ("scheme", "topo", "friction", "head_loss", "neg_depth_detection", "h_min",
              "phi_min", "cfl_max", "dt_max", "dt_map", "t0", "t_max", "v_max", "fr_max"):
        setattr(c, f, getattr(p, f))
    return c, keep


def _e_decl(math="glibc"):
    L = lib(math)
    if getattr(L, "_e_ready", False):
        return L
    L.orc_e_record.argtypes = [C.POINTER(OrcCase), _d, _d, _d, _d, _d, _d, C.c_int32, _d, C.c_int64,
                               _d, _d, _i32]
    L.orc_e_record.restype = C.c_int64
    L.orc_e_record_states.argtypes = [C.POINTER(OrcCase), _d, _d, _d, _d, _d, _d, C.c_int32, _d,
                                      C.c_int64, _d, _d, _i32, _d]
    L.orc_e_record_states.restype = C.c_int64
    L.orc_e_replay.argtypes = [C.POINTER(OrcCase), _d, _d, _d, _d, _d, _d, C.c_int64, _d, _d, _i32,
                               C.c_int32, C.c_int32, _i64, _d, _d, C.c_int32, _d, _d, _d]
    L.orc_e_replay.restype = C.c_double
    L.orc_e_gradient.argtypes = [C.POINTER(OrcCase), _d, _d, _d, _d, _d, _d, C.c_int64, _d, _d,
                                 _i32, C.c_int32, C.c_int32, _i64, _d, _d, C.c_int32, _d, _d, _d,
                                 _d]
    L.orc_e_gradient.restype = C.c_double
    L.orc_e_replay_frames.argtypes = [C.POINTER(OrcCase), _d, _d, _d, _d, _d, _d, C.c_int64, _d, _d,
                                      _i32, C.c_int32, _d, _d, _d, C.c_int32]
    L.orc_e_replay_frames.restype = C.c_double
    L.orc_e_gradient_frames.argtypes = [C.POINTER(OrcCase), _d, _d, _d, _d, _d, _d, C.c_int64, _d,
                                        _d, _i32, C.c_int32, _d, _d, _d, C.c_int32, _d, _d, _d, _d]
    L.orc_e_gradient_frames.restype = C.c_double
    L.orc_e_set_adjoint_cut.argtypes = [C.c_double]
    L.orc_e_set_adjoint_cut.restype = None
    L.orc_e_set_explosion_guard.argtypes = [C.c_double, C.c_int64]
    L.orc_e_set_explosion_guard.restype = None
    L.orc_e_truncation.argtypes = []
    L.orc_e_truncation.restype = C.c_int64
    L.orc_e_set_plant.argtypes = [C.c_int]
    L.orc_e_set_plant.restype = None
    L.orc_e_set_checkpoints.argtypes = [C.c_int64, _d]
    L.orc_e_set_checkpoints.restype = None
    L.orc_e_dry_states.argtypes = [C.POINTER(OrcCase), _d, _d, _d, _d, _d, _d, C.c_int64, _d, _d, _d]
    L.orc_e_dry_states.restype = None
    L._e_ready = True
    return L


class OracleE:
    """CPU restatement of the reference's _E forward: dry run and frozen replay."""

    def __init__(self, case, math="glibc"):
        self.case = case
        self.math = math
        self.c, self._keep = build_case(case)
        self.nc = case.n_cell
        self.ns = max(case.n_series, 1)

    def _ctl(self, x, default):
        v = default if x is None else np.broadcast_to(np.asarray(x, dtype=np.float64), (self.nc,))
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._keep.append(v)
        return v

    def record(self, obs_t, phi=None, K1=None, K2=None, max_steps=100000, states=False):
        a = self.case.a
        phi = self._ctl(phi, a["phi"])
        K1 = self._ctl(K1, a["fc1"] if len(a["fc1"]) else np.zeros(self.nc))
        K2 = self._ctl(K2, a["fc2"] if len(a["fc2"]) else np.zeros(self.nc))
        obs_t = np.ascontiguousarray(obs_t, dtype=np.float64)
        dt = np.empty(max_steps)
        bc = np.empty(max_steps * self.ns)
        trig = np.empty(max_steps, dtype=np.int32)
        h0, qx0, qy0 = (np.ascontiguousarray(a[k]) for k in ("h0", "qx0", "qy0"))
        S = np.empty((max_steps + 1) * 3 * self.nc) if states else None
        n = _e_decl(self.math).orc_e_record_states(C.byref(self.c), _dp(phi), _dp(K1), _dp(K2), _dp(h0),
                                          _dp(qx0), _dp(qy0), len(obs_t), _dp(obs_t), max_steps,
                                          _dp(dt), _dp(bc), trig.ctypes.data_as(_i32),
                                          _dp(S) if states else None)
        rec = {"dt": dt[:n].copy(), "bc": bc[: n * self.ns].reshape(n, self.ns).copy(),
               "trig": trig[:n].copy()}
        if states:  # the dry run's states [step][h | qx | qy]
            rec["states"] = S[: (n + 1) * 3 * self.nc].reshape(n + 1, 3, self.nc).copy()
        return rec

    def replay(self, rec, phi=None, K1=None, K2=None, gauges=None, obs=None, normalize=False,
               n_obs=None, bc=None):
        a = self.case.a
        phi = self._ctl(phi, a["phi"])
        K1 = self._ctl(K1, a["fc1"] if len(a["fc1"]) else np.zeros(self.nc))
        K2 = self._ctl(K2, a["fc2"] if len(a["fc2"]) else np.zeros(self.nc))
        dt = np.ascontiguousarray(rec["dt"], dtype=np.float64)
        bcs = np.ascontiguousarray(rec["bc"] if bc is None else bc, dtype=np.float64).reshape(-1)
        trig = np.ascontiguousarray(rec["trig"], dtype=np.int32)
        h0, qx0, qy0 = (np.ascontiguousarray(a[k]) for k in ("h0", "qx0", "qy0"))
        gc = np.ascontiguousarray(gauges if gauges is not None else np.zeros(0), dtype=np.int64)
        gw = np.ones(len(gc))
        ob = np.ascontiguousarray(obs if obs is not None else np.zeros(1), dtype=np.float64).reshape(-1)
        nob = int(n_obs if n_obs is not None else (trig.max() + 1 if len(trig) else 0))
        H, QX, QY = np.empty(self.nc), np.empty(self.nc), np.empty(self.nc)
        J = _e_decl(self.math).orc_e_replay(C.byref(self.c), _dp(phi), _dp(K1), _dp(K2), _dp(h0), _dp(qx0),
                                   _dp(qy0), len(dt), _dp(dt), _dp(bcs), trig.ctypes.data_as(_i32),
                                   nob, len(gc), gc.ctypes.data_as(_i64), _dp(gw), _dp(ob),
                                   1 if normalize else 0, _dp(H), _dp(QX), _dp(QY))
        return J, (H, QX, QY)


def _oracle_e_gradient(self, rec, phi=None, K1=None, K2=None, gauges=None, obs=None,
                       normalize=False, n_obs=None, bc=None, gauge_w=None, adj_cut=0.0,
                       explosion=None, checkpoints=None):
    """Gauge cost of the frozen replay and its gradient (reverse sweep of the
    oracle's own replay): {"J", "phi", "K1", "K2", "bc" [n_steps, n_series],
    "first_kept_step"}.  adj_cut > 0: thin-film cut of the reverse sweep
    (orc_e_set_adjoint_cut); explosion = (ratio, stride): the explosion guard
    over segments of `stride` steps (orc_e_set_explosion_guard); checkpoints =
    (stride, states): segments replayed from the dry run's states
    [n_steps + 1, 3, n_cell] (orc_e_set_checkpoints)."""
    a = self.case.a
    phi = self._ctl(phi, a["phi"])
    K1 = self._ctl(K1, a["fc1"] if len(a["fc1"]) else np.zeros(self.nc))
    K2 = self._ctl(K2, a["fc2"] if len(a["fc2"]) else np.zeros(self.nc))
    dt = np.ascontiguousarray(rec["dt"], dtype=np.float64)
    n = len(dt)
    bcs = np.ascontiguousarray(rec["bc"] if bc is None else bc, dtype=np.float64).reshape(-1)
    trig = np.ascontiguousarray(rec["trig"], dtype=np.int32)
    h0, qx0, qy0 = (np.ascontiguousarray(a[k]) for k in ("h0", "qx0", "qy0"))
    gc = np.ascontiguousarray(gauges if gauges is not None else np.zeros(0), dtype=np.int64)
    gw = np.ascontiguousarray(np.ones(len(gc)) if gauge_w is None else gauge_w, dtype=np.float64)
    ob = np.ascontiguousarray(obs if obs is not None else np.zeros(1), dtype=np.float64).reshape(-1)
    nob = int(n_obs if n_obs is not None else (trig.max() + 1 if len(trig) else 0))
    gphi, gK1, gK2 = np.empty(self.nc), np.empty(self.nc), np.empty(self.nc)
    gbc = np.empty(max(n, 1) * self.ns)
    L = _e_decl(self.math)
    L.orc_e_set_adjoint_cut(float(adj_cut))
    if explosion is not None:
        L.orc_e_set_explosion_guard(float(explosion[0]), int(explosion[1]))
    if checkpoints is not None:
        ck = np.ascontiguousarray(checkpoints[1], dtype=np.float64)
        if ck.size != (n + 1) * 3 * self.nc:
            raise ValueError("checkpoint states must be [n_steps + 1, 3, n_cell]")
        L.orc_e_set_checkpoints(int(checkpoints[0]), _dp(ck))
    try:
        J = L.orc_e_gradient(C.byref(self.c), _dp(phi), _dp(K1), _dp(K2), _dp(h0), _dp(qx0),
                             _dp(qy0), n, _dp(dt), _dp(bcs), trig.ctypes.data_as(_i32), nob,
                             len(gc), gc.ctypes.data_as(_i64), _dp(gw), _dp(ob),
                             1 if normalize else 0, _dp(gphi), _dp(gK1), _dp(gK2), _dp(gbc))
        kept = int(L.orc_e_truncation())
    finally:
        L.orc_e_set_adjoint_cut(0.0)
        L.orc_e_set_explosion_guard(0.0, 0)
        L.orc_e_set_checkpoints(0, None)
    return {"J": J, "phi": gphi, "K1": gK1, "K2": gK2, "bc": gbc[: n * self.ns].reshape(n, self.ns),
            "first_kept_step": kept}


OracleE.gradient = _oracle_e_gradient


def _oracle_e_dry_states(self, rec, phi=None, K1=None, K2=None, bc=None):
    """The dry run's states on the schedule rec (reconstruction with the
    previous step's dt): [n_steps + 1, 3, n_cell] (orc_e_dry_states)."""
    a = self.case.a
    phi = self._ctl(phi, a["phi"])
    K1 = self._ctl(K1, a["fc1"] if len(a["fc1"]) else np.zeros(self.nc))
    K2 = self._ctl(K2, a["fc2"] if len(a["fc2"]) else np.zeros(self.nc))
    dt = np.ascontiguousarray(rec["dt"], dtype=np.float64)
    bcs = np.ascontiguousarray(rec["bc"] if bc is None else bc, dtype=np.float64).reshape(-1)
    h0, qx0, qy0 = (np.ascontiguousarray(a[k]) for k in ("h0", "qx0", "qy0"))
    S = np.empty((len(dt) + 1) * 3 * self.nc)
    _e_decl(self.math).orc_e_dry_states(C.byref(self.c), _dp(phi), _dp(K1), _dp(K2), _dp(h0), _dp(qx0),
                               _dp(qy0), len(dt), _dp(dt), _dp(bcs), _dp(S))
    return S.reshape(len(dt) + 1, 3, self.nc)


OracleE.dry_states = _oracle_e_dry_states


def _oracle_e_frame_cost(self, rec, frames, phi=None, K1=None, K2=None, w_h=None, w_q=None,
                         normalize=False, n_obs=None):
    """Frame-misfit cost of the frozen replay; frames [n_obs, 3, n_cell] or
    [n_obs, n_cell, 3] (h, qx, qy)."""
    a = self.case.a
    phi = self._ctl(phi, a["phi"])
    K1 = self._ctl(K1, a["fc1"] if len(a["fc1"]) else np.zeros(self.nc))
    K2 = self._ctl(K2, a["fc2"] if len(a["fc2"]) else np.zeros(self.nc))
    fr = np.asarray(frames, dtype=np.float64)
    if fr.shape[-1] == 3 and fr.shape[1] == self.nc:
        fr = fr.transpose(0, 2, 1)
    fr = np.ascontiguousarray(fr)
    self._keep.append(fr)
    dt = np.ascontiguousarray(rec["dt"], dtype=np.float64)
    bcs = np.ascontiguousarray(rec["bc"], dtype=np.float64).reshape(-1)
    trig = np.ascontiguousarray(rec["trig"], dtype=np.int32)
    h0, qx0, qy0 = (np.ascontiguousarray(a[k]) for k in ("h0", "qx0", "qy0"))
    wh = None if w_h is None else _dp(self._ctl(w_h, None))
    wq = None if w_q is None else _dp(self._ctl(w_q, None))
    nob = int(n_obs if n_obs is not None else fr.shape[0])
    return _e_decl(self.math).orc_e_replay_frames(C.byref(self.c), _dp(phi), _dp(K1), _dp(K2), _dp(h0),
                                         _dp(qx0), _dp(qy0), len(dt), _dp(dt), _dp(bcs),
                                         trig.ctypes.data_as(_i32), nob, _dp(fr), wh, wq,
                                         1 if normalize else 0)


OracleE.frame_cost = _oracle_e_frame_cost


def _oracle_e_frame_gradient(self, rec, frames, phi=None, K1=None, K2=None, normalize=False,
                             n_obs=None):
    """Frame-misfit cost and its gradient (uniform weights)."""
    a = self.case.a
    phi = self._ctl(phi, a["phi"])
    K1 = self._ctl(K1, a["fc1"] if len(a["fc1"]) else np.zeros(self.nc))
    K2 = self._ctl(K2, a["fc2"] if len(a["fc2"]) else np.zeros(self.nc))
    fr = np.asarray(frames, dtype=np.float64)
    if fr.shape[-1] == 3 and fr.shape[1] == self.nc:
        fr = fr.transpose(0, 2, 1)
    fr = np.ascontiguousarray(fr)
    self._keep.append(fr)
    dt = np.ascontiguousarray(rec["dt"], dtype=np.float64)
    n = len(dt)
    bcs = np.ascontiguousarray(rec["bc"], dtype=np.float64).reshape(-1)
    trig = np.ascontiguousarray(rec["trig"], dtype=np.int32)
    h0, qx0, qy0 = (np.ascontiguousarray(a[k]) for k in ("h0", "qx0", "qy0"))
    gphi, gK1, gK2 = np.empty(self.nc), np.empty(self.nc), np.empty(self.nc)
    gbc = np.empty(max(n, 1) * self.ns)
    nob = int(n_obs if n_obs is not None else fr.shape[0])
    J = _e_decl(self.math).orc_e_gradient_frames(C.byref(self.c), _dp(phi), _dp(K1), _dp(K2), _dp(h0),
                                        _dp(qx0), _dp(qy0), n, _dp(dt), _dp(bcs),
                                        trig.ctypes.data_as(_i32), nob, _dp(fr), None, None,
                                        1 if normalize else 0, _dp(gphi), _dp(gK1), _dp(gK2),
                                        _dp(gbc))
    return {"J": J, "phi": gphi, "K1": gK1, "K2": gK2, "bc": gbc[: n * self.ns].reshape(n, self.ns)}


OracleE.frame_gradient = _oracle_e_frame_gradient


class OracleSolver:
    """CPU restatement of swSp's time loop on one case (arrays in reference order)."""

    def __init__(self, case, params=None, math="glibc"):
        a = case.a
        p = params if params is not None else case.params
        self.math = math
        self.n_cell = case.n_cell
        self._keep = []

        def k(arr, dt=np.float64):
            arr = np.ascontiguousarray(arr, dtype=dt)
            self._keep.append(arr)
            return arr

        c = OrcCase()
        c.n_cell, c.n_bnd, c.n_int = case.n_cell, case.n_bnd, case.n_int
        for f in ("cell_area", "cell_x", "cell_y", "cell_zb", "phi", "beta", "bnd_len", "bnd_nx",
                  "bnd_ny", "int_len", "int_nx", "int_ny", "int_xm", "int_ym"):
            setattr(c, f, _dp(k(a[f])))
        for f, src in (("fric_c1", "fc1"), ("fric_c2", "fc2"), ("fric_alpha_deg", "falpha"),
                       ("hl_hsing", "hsing"), ("hl_c1", "hk1"), ("hl_c2", "hk2"),
                       ("hl_alpha_deg", "halpha")):
            setattr(c, f, _dp(k(a[src])))
        c.bnd_cell = k(a["bnd_cell"], np.int64).ctypes.data_as(_i64)
        c.bnd_series = k(a["bnd_series"], np.int32).ctypes.data_as(_i32)
        c.int_L = k(a["int_L"], np.int64).ctypes.data_as(_i64)
        c.int_R = k(a["int_R"], np.int64).ctypes.data_as(_i64)
        c.n_series, c.n_times = case.n_series, case.n_times
        c.series_type = k(a["series_type"], np.int32).ctypes.data_as(_i32)
        c.series_t = _dp(k(a["series_t"]))
        c.series_v = _dp(k(a["series_v"]))
        for f in ("scheme", "topo", "friction", "head_loss", "neg_depth_detection", "h_min",
                  "phi_min", "cfl_max", "dt_max", "dt_map", "t0", "t_max", "v_max", "fr_max"):
            setattr(c, f, getattr(p, f))
        wc = a["wind_code"] if "wind_code" in a else np.zeros(0, np.int32)
        if len(wc) > 0:
            c.n_wind_times = len(a["wind_t"])
            c.n_wind_station = len(a["wind_x"]) // c.n_wind_times
            c.wind_code = k(wc, np.int32).ctypes.data_as(_i32)
            c.wind_t, c.wind_x, c.wind_y = (_dp(k(a[f])) for f in ("wind_t", "wind_x", "wind_y"))
        models = a["source_models"] if "source_models" in a else np.zeros(3, np.int32)
        c.runoff_model, c.rain_water_only, c.infil_model = (int(x) for x in models[:3])
        c.momentum_diffusion = int(models[3]) if len(models) > 3 else 0
        if c.runoff_model:
            c.n_rain_times = len(a["rain_t"])
            c.n_rain_station = len(a["rain_v"]) // c.n_rain_times
            c.rain_code = k(a["rain_code"], np.int32).ctypes.data_as(_i32)
            c.rain_t, c.rain_v = _dp(k(a["rain_t"])), _dp(k(a["rain_v"]))
            if c.runoff_model != 1:
                c.runoff_coeff = _dp(k(a["runoff_coeff"]))
        if c.infil_model:
            c.infil_rate = _dp(k(a["infil_rate"]))
        h0, qx0, qy0 = k(a["h0"]), k(a["qx0"]), k(a["qy0"])
        self._s = lib(self.math).orc_create(C.byref(c), _dp(h0), _dp(qx0), _dp(qy0))
        if not self._s:
            raise RuntimeError("orc_create failed")

    def __del__(self):
        if getattr(self, "_s", None):
            lib(self.math).orc_destroy(self._s)
            self._s = None

    def step(self, n=1):
        done = C.c_int64()
        st = lib(self.math).orc_run(self._s, 1e300, int(n), C.byref(done))
        if st:
            raise RuntimeError(f"oracle error {st}")
        return done.value

    def run(self, t_end=1e300, max_steps=2 ** 62):
        done = C.c_int64()
        st = lib(self.math).orc_run(self._s, float(t_end), int(max_steps), C.byref(done))
        if st:
            raise RuntimeError(f"oracle error {st}")
        return done.value

    def state(self):
        n = self.n_cell
        h, qx, qy = np.empty(n), np.empty(n), np.empty(n)
        lib(self.math).orc_get_state(self._s, _dp(h), _dp(qx), _dp(qy))
        return h, qx, qy

    @property
    def t(self):
        return lib(self.math).orc_time(self._s)

    @property
    def dt(self):
        return lib(self.math).orc_last_dt(self._s)


def run_reference(folder, timeout=600):
    """Run the compiled reference sw2dSolver (oracle/_ref) on a case folder."""
    if not os.path.exists(REF_SOLVER):
        raise FileNotFoundError(REF_SOLVER)
    r = subprocess.run([REF_SOLVER, str(folder)], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"reference solver failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
    return r
