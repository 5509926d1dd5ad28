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

``minimize`` is L-BFGS-B (memory 6, at most 30 line-search evaluations, the
reference's settings for its functional case) over box bounds, run by the
native minimiser of the host library.
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


LBFGSB_STOP = ("converged (projected gradient)", "converged (objective change)",
               "max iterations", "line search failed", "stopped by the caller")


def lbfgsb(fg, x0, lo=None, hi=None, m=6, eps=1e-5, eps_rel=None, max_iter=0, max_ls=20, past=1,
           delta=1e-10, ftol=1e-4, wolfe=0.9, first_step="unit", max_submin=10, restart=False,
           stop=None):
    """L-BFGS-B over fg(x) -> (value, gradient): the native minimiser
    (sw2dh_lbfgsb, include/sw2d_host.h), which stands in for the reference's
    LBFGSpp::LBFGSBSolver (assim/swSpE_minimize.cpp:857-898; parameters m =
    lbfgsb_m, epsilon = epsilon_rel = lbfgsb_epsilon, max_linesearch =
    lbfgsb_max_linesearch, max_iterations = max_it).  Every evaluation is a
    value and a gradient.  first_step "unit" (LBFGSpp) tries the first
    direction at unit length, "cauchy" (the 1997 Fortran code on boxed
    problems) goes to the first Cauchy point.  max_submin bounds the passes of
    the box-constrained subspace minimisation.  restart=True drops the
    correction pairs after a failed line search and searches again from the
    last iterate (the Fortran L-BFGS-B 3.0 rule); False stops there (LBFGSpp,
    whose exception the reference catches, keeping the last iterate).
    stop(x, value) -> True ends the run after that
    evaluation (its point is kept when it is the best).  Returns (x, value,
    info) with info = {n_iter, n_eval, pgnorm, reason, trace [(wall s, value,
    'grad')]}."""
    import ctypes as C

    from . import _native as N
    lib = N.host_lib()
    x = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).copy())
    n = x.size
    lo_a = None if lo is None else np.ascontiguousarray(np.broadcast_to(np.asarray(lo, float), x.shape))
    hi_a = None if hi is None else np.ascontiguousarray(np.broadcast_to(np.asarray(hi, float), x.shape))
    P = N.LbfgsbParam()
    lib.sw2dh_lbfgsb_default(C.byref(P))
    P.m, P.epsilon = int(m), float(eps)
    P.epsilon_rel = float(eps if eps_rel is None else eps_rel)
    P.past, P.delta = int(past), float(delta)
    P.max_iterations, P.max_linesearch = int(max_iter), int(max_ls)
    P.ftol, P.wolfe = float(ftol), float(wolfe)
    if first_step not in ("unit", "cauchy"):
        raise ValueError("first_step must be 'unit' or 'cauchy'")
    P.first_step = 0 if first_step == "unit" else 1
    P.max_submin, P.restart = int(max_submin), int(bool(restart))
    t0 = time.perf_counter()
    trace, err = [], []

    def cb(xp, fp, gp, _user):
        try:
            xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
            J, g = fg(xv)
            J = float(J)
            g = np.asarray(g, dtype=np.float64).reshape(n)
            fp[0] = J
            np.ctypeslib.as_array(gp, shape=(n,))[:] = g
            trace.append((time.perf_counter() - t0, J, "grad"))
            return 1 if (stop is not None and stop(xv, J)) else 0
        except BaseException as e:  # noqa: BLE001 -- re-raised below
            err.append(e)
            return -1

    cfun = N.OBJECTIVE(cb)
    fx, pg = C.c_double(0.0), C.c_double(0.0)
    it, ne, why = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    rc = lib.sw2dh_lbfgsb(n, N.dptr(x), N.dptr(lo_a), N.dptr(hi_a), cfun, None, C.byref(P),
                          C.byref(fx), C.byref(it), C.byref(ne), C.byref(pg), C.byref(why))
    if err:
        raise err[0]
    if rc != 0:
        raise RuntimeError(lib.sw2dh_lbfgsb_error().decode())
    return x, fx.value, dict(n_iter=it.value, n_eval=ne.value, pgnorm=pg.value,
                             reason=LBFGSB_STOP[why.value], trace=trace)


def minimize(fg, f=None, x0=None, lo=None, hi=None, m=6, max_iter=100, max_ls=30, eps=1e-5,
             time_limit=None, stop=None, restart=False):
    """run_minimization's "lbfgsb" branch over box bounds: L-BFGS-B (memory m,
    at most max_ls line-search evaluations, projected-gradient tolerance eps
    absolute and relative, the reference's settings for its functional case),
    every evaluation a gradient fg(x) -> (value, gradient) as the reference's
    objective adaptor (swSpE_minimize.cpp:850-856).  f is not used (kept for
    callers that pass a cost-only function).  Stops also after time_limit
    seconds or when stop(x, value) is true; restart as in lbfgsb.  Returns
    (x, value, trace)."""
    t0 = time.perf_counter()

    def halt(xv, J):
        if time_limit and time.perf_counter() - t0 > time_limit:
            return True
        return bool(stop(xv, J)) if stop is not None else False

    x, J, info = lbfgsb(fg, x0, lo=lo, hi=hi, m=m, eps=eps, max_iter=max_iter, max_ls=max_ls,
                        stop=halt, restart=restart)
    return x, J, info["trace"]


def minimize_gd(fg, f, x0, project=True, step=1e-6, max_iter=100, c1=1e-4, shrink=0.5, grow=4.0,
                max_bt=40):
    """Projected gradient descent with Armijo backtracking: run_minimization's
    "gd" (assim/swSpE_minimize.cpp:747-831).  Each iterate takes one gradient
    fg(x) -> (J, g); the trial step starts at the last accepted step times
    `grow` (the first at `step`), is multiplied by `shrink` up to `max_bt`
    times until f(x_trial) is finite and <= J - c1 g.(x - x_trial) with
    g.(x - x_trial) > 0, x_trial = x - t g clipped to [0, 1] when `project`.
    Stops at max_iter, on a zero / NaN gradient or when no step is accepted.
    Returns the best iterate by gradient-evaluated J: (x, J, trace)."""
    x = np.asarray(x0, dtype=np.float64).copy()
    t = step if step > 0.0 else 1e-6
    best_x, best_J = x.copy(), np.inf
    t0 = time.perf_counter()
    trace = []
    for _ in range(max_iter):
        J, g = fg(x)
        trace.append((time.perf_counter() - t0, J, "grad"))
        if J < best_J:
            best_J, best_x = J, x.copy()
        if not np.dot(g, g) > 0.0:
            break
        t_try = t * grow
        accepted = False
        for _ in range(max_bt):
            v = x - t_try * g
            if project:
                v = np.clip(v, 0.0, 1.0)
            pg = float(np.dot(g, x - v))
            Jn = f(v)
            trace.append((time.perf_counter() - t0, Jn, "ls"))
            if np.isfinite(Jn) and Jn <= J - c1 * pg and pg > 0.0:
                accepted = True
                break
            t_try *= shrink
        if not accepted:
            break
        x, t = v, t_try
    return best_x, best_J, trace


def minimize_adam(fg, x0, step=1e-2, max_iter=100, project=True, safe=False, b1=0.9, b2=0.999,
                  eps=1e-8, explosion=50.0):
    """Projected Adam: run_minimization's "adam" (assim/swSpE_minimize.cpp:
    932-1017).  One gradient per iterate; m, v moments with bias correction,
    x -= step mhat / (sqrt(vhat) + eps), clipped to [0, 1] when `project`.
    A jump of ||g||_inf above `explosion` x the previous one stops the loop
    when `safe` (otherwise it is only reported in the trace).  Returns the
    best iterate by J: (x, J, trace)."""
    x = np.asarray(x0, dtype=np.float64).copy()
    m = np.zeros_like(x)
    v = np.zeros_like(x)
    best_x, best_J = x.copy(), np.inf
    g_prev = -1.0
    t0 = time.perf_counter()
    trace = []
    for it in range(1, max_iter + 1):
        J, g = fg(x)
        trace.append((time.perf_counter() - t0, J, "grad"))
        if J < best_J:
            best_J, best_x = J, x.copy()
        g_inf = float(np.max(np.abs(g))) if len(g) else 0.0
        if g_prev > 0.0 and g_inf > explosion * g_prev:
            trace.append((time.perf_counter() - t0, J, "explosion"))
            if safe:
                break
        g_prev = g_inf
        bc1 = 1.0 - b1 ** it
        bc2 = 1.0 - b2 ** it
        m = b1 * m + (1.0 - b1) * g
        v = b2 * v + (1.0 - b2) * g * g
        x = x - step * (m / bc1) / (np.sqrt(v / bc2) + eps)
        if project:
            x = np.clip(x, 0.0, 1.0)
    return best_x, best_J, trace


def _wolfe(fp, x, d, J0, g0, c1, c2, t=1.0, max_eval=40):
    """Step along d satisfying the strong Wolfe conditions: bracketing then
    zoom by bisection (Nocedal & Wright, Numerical Optimization, algorithms
    3.5 and 3.6).  fp(x) -> (J, g).  Returns (t, J, g), or None after
    max_eval evaluations."""
    d0 = float(np.dot(g0, d))
    n = [0]

    def ev(s):
        n[0] += 1
        J, g = fp(x + s * d)
        return J, g, float(np.dot(g, d))

    def zoom(lo, J_lo, hi):
        while n[0] < max_eval:
            s = 0.5 * (lo + hi)
            J, g, ds = ev(s)
            if not np.isfinite(J) or J > J0 + c1 * s * d0 or J >= J_lo:
                hi = s
            else:
                if abs(ds) <= -c2 * d0:
                    return s, J, g
                if ds * (hi - lo) >= 0.0:
                    hi = lo
                lo, J_lo = s, J
        return None

    t_prev, J_prev = 0.0, J0
    while n[0] < max_eval:
        J, g, ds = ev(t)
        if not np.isfinite(J) or J > J0 + c1 * t * d0 or (t_prev > 0.0 and J >= J_prev):
            return zoom(t_prev, J_prev, t)
        if abs(ds) <= -c2 * d0:
            return t, J, g
        if ds >= 0.0:
            return zoom(t, J, t_prev)
        t_prev, J_prev, t = t, J, 2.0 * t
    return None


def minimize_bfgs(fg, x0, max_iter=100, rel_eps=1e-8, c1=1e-4, c2=0.9):
    """Dense BFGS on the inverse Hessian with a strong-Wolfe line search, the
    method of run_minimization's default branch (PSQN::bfgs with rel_eps 1e-8,
    c1 1e-4, c2 0.9, assim/swSpE_minimize.cpp:913-930; the PSQN library is
    not part of the reference tree, so this follows the textbook method:
    Nocedal & Wright, algorithm 6.1, H0 scaled by s.y / y.y after the first
    step).  Stops when |J_new - J| <= rel_eps (|J| + rel_eps), at max_iter, or
    when the line search fails.  Returns (x, J, trace)."""
    x = np.asarray(x0, dtype=np.float64).copy()
    n = len(x)
    t0 = time.perf_counter()
    J, g = fg(x)
    trace = [(time.perf_counter() - t0, J, "grad")]
    H = np.eye(n)
    first = True
    for _ in range(max_iter):
        d = -H @ g
        if not np.dot(d, g) < 0.0:
            H = np.eye(n)
            d = -g
        res = _wolfe(fg, x, d, J, g, c1, c2)
        if res is None:
            break
        t, Jn, gn = res
        trace.append((time.perf_counter() - t0, Jn, "grad"))
        s, y = t * d, gn - g
        sy = float(np.dot(s, y))
        if sy > 0.0:
            if first:
                H = np.eye(n) * (sy / float(np.dot(y, y)))
                first = False
            rho = 1.0 / sy
            Hy = H @ y
            H = H - rho * (np.outer(s, Hy) + np.outer(Hy, s)) + (rho * rho * float(np.dot(y, Hy)) + rho) * np.outer(s, s)
        done = abs(Jn - J) <= rel_eps * (abs(J) + rel_eps)
        x, J, g = x + s, Jn, gn
        if done:
            break
    return x, J, trace


METHODS = ("lbfgsb", "gd", "adam", "bfgs")


def run_minimization(prob, method="lbfgsb", max_iter=100, step=None, project=None, **kw):
    """The reference's run_minimization dispatch over PorosityProblem
    (assim/swSpE_minimize.cpp:725-1017): "lbfgsb" (L-BFGS-B on [0, 1]
    with memory 6 and <= 30 line-search evaluations), "gd" (projected GD +
    Armijo), "adam" (projected Adam), "bfgs" (dense BFGS, unconstrained).
    Projection onto [0, 1] applies, as in the reference, when the change of
    variable leaves m = phi (none / clamp).  Returns (theta, J, trace)."""
    if method not in METHODS:
        raise ValueError(f"method must be one of {METHODS}")
    if project is None:
        project = prob.mode in (None, "none", "clamp")
    theta0 = prob.theta0()
    if method == "lbfgsb":
        lo, hi = (0.0, 1.0) if project else (None, None)
        # a probe gradient at the start fixes where the adjoint truncates; the
        # window stays frozen there for the whole solve so that every
        # line-search evaluation covers the same segments
        # (swSpE_minimize.cpp:881-903)
        A = getattr(prob, "A", None)
        if A is None or not hasattr(A, "set_first_segment"):
            return minimize(prob.grad, prob.func, theta0, lo=lo, hi=hi, max_iter=max_iter, **kw)
        prob.grad(theta0)
        A.set_first_segment(max(0, A.truncation()[1]))
        try:
            return minimize(prob.grad, prob.func, theta0, lo=lo, hi=hi, max_iter=max_iter, **kw)
        finally:
            A.set_first_segment(None)
    if method == "gd":
        return minimize_gd(prob.grad, prob.func, theta0, project=project, max_iter=max_iter,
                           step=1e-6 if step is None else step, **kw)
    if method == "adam":
        return minimize_adam(prob.grad, theta0, step=1e-2 if step is None else step,
                             max_iter=max_iter, project=project, **kw)
    return minimize_bfgs(prob.grad, theta0, max_iter=max_iter, **kw)


def qsva_window_ends(t0, t_max, delta_tau):
    """Window ends t0 + k delta_tau (k = 1, 2, ...) below t_max, then t_max."""
    if not delta_tau > 0.0 or not t_max > t0:
        raise ValueError("qsva: need delta_tau > 0 and t_max > t0")
    ends, k = [], 1
    while t0 + k * delta_tau < t_max:
        ends.append(t0 + k * delta_tau)
        k += 1
    return ends + [t_max]


def run_qsva(make_problem, t0, t_max, delta_tau, phi_start, method="lbfgsb", max_iter=100,
             on_step=None, **kw):
    """Quasi-static variational assimilation: the reference documents it
    (Doc_assim_4Dvar/txt/Documentation_détaillé_assim.txt:1101-1160, swSp::run_qsva,
    keys qsva_enabled / qsva_delta_tau) but does not implement it.  The window
    start t0 stays fixed and its end grows by delta_tau up to t_max; each
    window's minimisation (run_minimization, the same methods) starts from the
    previous window's optimum.  make_problem(t_end, phi) returns the
    PorosityProblem of the window [t0, t_end] (its observations in that window)
    whose starting point is phi; phi_start starts the first window.  Returns
    (phi of the full window, per-window records {t_end, J, n_eval, phi})."""
    phi = np.asarray(phi_start, dtype=np.float64).copy()
    steps = []
    for t_end in qsva_window_ends(float(t0), float(t_max), float(delta_tau)):
        prob = make_problem(t_end, phi)
        theta, J, trace = run_minimization(prob, method, max_iter=max_iter, **kw)
        phi = prob.m_of(theta)
        rec = dict(t_end=t_end, J=float(J), n_eval=len(trace), phi=phi.copy())
        steps.append(rec)
        if on_step is not None:
            on_step(rec)
    return phi, steps
