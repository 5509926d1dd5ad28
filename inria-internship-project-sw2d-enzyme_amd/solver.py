"""GPU single-porosity solver: the host-side mirror of the reference's swSp.

``SpSolver(case)`` plays the part of ``swParent::create("sp")`` +
``swSp::init`` (src/sw2dSolver/models/sp/swSp.cpp:720-969); ``execute()`` is
``swSp::execute`` (swSp.cpp:482-670); ``step(n)`` runs n iterations of its
time loop.  All arithmetic runs in the HIP kernels of libsw2dg.so; there is
no CPU path here.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .cases import Case, MemCase


class SpSolver:
    def __init__(self, case, device=-1, params=None):
        if not isinstance(case, (Case, MemCase)):
            case = Case(case)
        self.case = case
        self.lib = N.gpu_lib()
        a = case.a
        self._keep = {}

        def keep(name, arr, dtype):
            arr = np.ascontiguousarray(arr, dtype=dtype)
            self._keep[name] = arr
            return arr

        m = N.Mesh()
        m.n_cell, m.n_bnd, m.n_int = case.n_cell, case.n_bnd, case.n_int
        for f in ("cell_area", "cell_x", "cell_y", "cell_zb", "bnd_len", "bnd_nx", "bnd_ny",
                  "int_len", "int_nx", "int_ny", "int_xm", "int_ym"):
            setattr(m, f, N.dptr(keep(f, a[f], np.float64)))
        m.bnd_cell = N.i64ptr(keep("bnd_cell", a["bnd_cell"], np.int64))
        m.int_L = N.i64ptr(keep("int_L", a["int_L"], np.int64))
        m.int_R = N.i64ptr(keep("int_R", a["int_R"], np.int64))
        m.bnd_series = N.i32ptr(keep("bnd_series", a["bnd_series"], np.int32))
        self.params = params if params is not None else case.params
        ctx = C.c_void_p()
        rc = self.lib.sw2dg_ctx_create(C.byref(m), C.byref(self.params), device, C.byref(ctx))
        if rc != 0:
            raise N.SW2DError(rc, "sw2dg_ctx_create failed")
        self.ctx = ctx
        fric = self.params.friction != N.FRIC_NONE
        hl = self.params.head_loss != 0
        arr = lambda k, on=True: N.dptr(keep(k, a[k], np.float64)) if on else None
        self._check(self.lib.sw2dg_set_cell_params(
            ctx, arr("phi"), arr("beta"), arr("fc1", fric), arr("fc2", fric), arr("falpha", fric),
            arr("hsing", hl), arr("hk1", hl), arr("hk2", hl), arr("halpha", hl)))
        st = keep("series_type", a["series_type"], np.int32)
        sv = keep("series_v", a["series_v"], np.float64)
        stt = keep("series_t", a["series_t"], np.float64)
        self._check(self.lib.sw2dg_set_bc_series(ctx, case.n_series, case.n_times, N.i32ptr(st),
                                                 N.dptr(stt), N.dptr(sv)))
        models = a["source_models"] if "source_models" in a else (0, 1, 0, 0)
        if len(models) > 3 and models[3]:
            self.set_momentum_diffusion(1)
        if models[0]:
            nt = len(a["rain_t"])
            self.set_runoff(int(models[0]), bool(models[1]), a["rain_code"], a["rain_t"],
                            np.reshape(a["rain_v"], (-1, nt)),
                            a["runoff_coeff"] if models[0] != 1 else None)
        if models[2]:
            self.set_infiltration(int(models[2]), a["infil_rate"])
        if "wind_code" in a and len(a["wind_code"]) > 0:
            nt = len(a["wind_t"])
            self.set_wind(a["wind_code"], a["wind_t"], np.reshape(a["wind_x"], (-1, nt)),
                          np.reshape(a["wind_y"], (-1, nt)))
        self.set_state(a["h0"], a["qx0"], a["qy0"], self.params.t0)

    # -- plumbing -----------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise N.SW2DError(rc, self.lib.sw2dg_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.lib.sw2dg_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bc_series(self, t, v, types=None):
        """Replace the boundary time series: t [n_times], v [n_series, n_times]
        (swParent::initBoundaryConditions for time series).  Call set_state
        afterwards before running the native time loop."""
        t = self._hold("bc_t", t)
        v = np.atleast_2d(np.asarray(v, dtype=np.float64))
        ns = v.shape[0]
        if types is None:
            types = self.case.a["series_type"][:ns]
        ty = np.ascontiguousarray(types, dtype=np.int32)
        self._keep["bc_type"] = ty
        vv = self._hold("bc_v", v.reshape(-1))
        self._check(self.lib.sw2dg_set_bc_series(self.ctx, ns, len(t), N.i32ptr(ty), N.dptr(t),
                                                 N.dptr(vv)))

    def set_runoff(self, model, water_only, code=None, t=None, rate=None, coeff=None):
        """Runoff (swSp::setRunoffParameters + the precipitation tables,
        swSp.cpp:1276-1320, swParent.cpp:516-569): model 0 none, 1
        full_rainfall, 2 infiltration_coefficient, 3 runoff_coefficient;
        water_only is model_porosity_rainfall water_only; code[n_cell]
        1-based rain station per cell, t[n_times], rate[n_station, n_times]
        (m/s), coeff[n_cell] for models 2 and 3."""
        if not model:
            self._check(self.lib.sw2dg_set_runoff(self.ctx, 0, int(bool(water_only)), 0, 0, None,
                                                  None, None, None))
            return
        code = np.ascontiguousarray(code, dtype=np.int32)
        self._keep["rain_code"] = code
        t = self._hold("rain_t", t)
        rate = self._hold("rain_v", np.reshape(rate, (-1, len(t))))
        cf = None if coeff is None else N.dptr(self._hold("runoff_coeff", coeff))
        if code.shape != (self.case.n_cell,):
            raise ValueError("set_runoff: code[n_cell] expected")
        self._check(self.lib.sw2dg_set_runoff(self.ctx, int(model), int(bool(water_only)),
                                              rate.shape[0], len(t), N.i32ptr(code), N.dptr(t),
                                              N.dptr(rate), cf))

    def set_infiltration(self, model, rate=None):
        """Horton infiltration (swSp::setInfiltrationParameters, swSp.cpp:1075-1115):
        model 1 with rate[n_cell] (m/s), 0 off."""
        r = None if not model else N.dptr(self._hold("infil_rate", rate))
        self._check(self.lib.sw2dg_set_infiltration(self.ctx, int(model), r))

    def set_momentum_diffusion(self, model):
        """Momentum diffusion (swSp::parabComputeExplicit, swSpParab.cpp:10-96):
        1 on (velocity_independent, coefficient 1e-6 as the reference's), 0 off."""
        self._check(self.lib.sw2dg_set_momentum_diffusion(self.ctx, int(model)))

    def set_wind(self, code, t, vx, vy):
        """Wind stations (swParent::setWindStationCodes / setWindVelocityTimeSeriesTables,
        swParent.cpp:571-619): code[n_cell] 1-based station per cell, t[n_times]
        record times, vx, vy[n_station, n_times] Cartesian velocities.  An
        empty code array turns the source off."""
        code = np.ascontiguousarray(code, dtype=np.int32)
        self._keep["wind_code"] = code
        if code.size == 0:
            self._check(self.lib.sw2dg_set_wind(self.ctx, 0, 0, None, None, None, None))
            return
        t = self._hold("wind_t", np.ascontiguousarray(t, dtype=np.float64))
        vx = self._hold("wind_x", np.ascontiguousarray(vx, dtype=np.float64).reshape(-1, len(t)))
        vy = self._hold("wind_y", np.ascontiguousarray(vy, dtype=np.float64).reshape(-1, len(t)))
        if code.shape != (self.case.n_cell,) or vx.shape != vy.shape:
            raise ValueError("set_wind: code[n_cell] and vx, vy[n_station, n_times] expected")
        self._check(self.lib.sw2dg_set_wind(self.ctx, vx.shape[0], len(t), N.i32ptr(code),
                                            N.dptr(t), N.dptr(vx), N.dptr(vy)))

    def _hold(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._keep[name] = arr
        return arr

    # -- state ---------------------------------------------------------------
    def set_state(self, h, qx, qy, t):
        f = lambda x: np.ascontiguousarray(np.broadcast_to(x, (self.case.n_cell,)), dtype=np.float64)
        h, qx, qy = f(h), f(qx), f(qy)
        self._check(self.lib.sw2dg_set_state(self.ctx, N.dptr(h), N.dptr(qx), N.dptr(qy), float(t)))

    def state(self):
        n = self.case.n_cell
        h, qx, qy = np.empty(n), np.empty(n), np.empty(n)
        self._check(self.lib.sw2dg_get_state(self.ctx, N.dptr(h), N.dptr(qx), N.dptr(qy)))
        return h, qx, qy

    def clock(self):
        t, dt = C.c_double(), C.c_double()
        steps = C.c_int64()
        flags = C.c_int32()
        self._check(self.lib.sw2dg_get_clock(self.ctx, C.byref(t), C.byref(dt), C.byref(steps),
                                             C.byref(flags)))
        return {"t": t.value, "dt": dt.value, "steps": steps.value, "store_flags": flags.value}

    @property
    def t(self):
        return self.clock()["t"]

    def bnd_f1(self):
        out = np.empty(self.case.n_bnd)
        self._check(self.lib.sw2dg_get_bnd_f1(self.ctx, N.dptr(out)))
        return out

    # -- time loop -------------------------------------------------------------
    def run(self, t_stop=np.inf, max_steps=2 ** 62, pause_on_store=False):
        n = C.c_int64()
        self._check(self.lib.sw2dg_run(self.ctx, float(min(t_stop, 1e300)), int(max_steps),
                                       1 if pause_on_store else 0, C.byref(n)))
        return n.value

    def step(self, n=1):
        return self.run(max_steps=n)

    def execute(self):
        """swSp::execute: run to t_max."""
        return self.run(t_stop=self.params.t_max)

    def profile(self, n_steps):
        avg = (C.c_double * 8)()
        names = (C.c_char_p * 8)()
        nk = C.c_int32()
        self._check(self.lib.sw2dg_profile_kernels(self.ctx, int(n_steps), 8, avg, names,
                                                   C.byref(nk)))
        return {names[i].decode(): avg[i] for i in range(nk.value)}

    def device_bytes(self):
        return self.lib.sw2dg_device_bytes(self.ctx)
