"""The native L-BFGS-B (sw2dh_lbfgsb), which stands in for the reference's
LBFGSpp::LBFGSBSolver (assim/swSpE_minimize.cpp:832-912).  LBFGSpp is not in
the reference tree, so its iterates are not a parity target; the method is
pinned by what the mathematics fixes: closed-form constrained minima, the
KKT conditions of random bound-constrained quadratics (checked against an
active-set solution computed independently here), Rosenbrock with and
without an active bound, and a quasi-Newton evaluation count (a projected
gradient method, or a BFGS update with a wrong sign, needs far more)."""
import numpy as np
import pytest


def _lb():
    from conftest import load_pkg
    load_pkg()
    from sw2d_amd.porosity import lbfgsb
    return lbfgsb


def rosen(x):
    f = np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1.0 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] = -400.0 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2.0 * (1.0 - x[:-1])
    g[1:] += 200.0 * (x[1:] - x[:-1] ** 2)
    return f, g


def qp_active_set(A, b, lo, hi):
    """Exact minimiser of 0.5 x'Ax - b'x on a box (A SPD) by enumerating
    the projected-Newton active set until the KKT conditions hold."""
    n = len(b)
    act = np.zeros(n, dtype=int)  # -1 lower, +1 upper, 0 free
    for _ in range(200):
        x = np.where(act < 0, lo, np.where(act > 0, hi, 0.0))
        F = act == 0
        if F.any():
            rhs = b[F] - A[np.ix_(F, ~F)] @ x[~F]
            x[F] = np.linalg.solve(A[np.ix_(F, F)], rhs)
        g = A @ x - b
        viol = (F & ((x < lo - 1e-12) | (x > hi + 1e-12)))
        wrong = (~F) & (((act < 0) & (g < -1e-12)) | ((act > 0) & (g > 1e-12)))
        if not viol.any() and not wrong.any():
            return x
        i = np.argmax(np.where(viol, np.maximum(lo - x, x - hi), 0.0)) if viol.any() else None
        if i is not None:
            act[i] = -1 if x[i] < lo[i] else 1
        else:
            j = np.argmax(np.where(wrong, np.abs(g), 0.0))
            act[j] = 0
    raise RuntimeError("active set did not settle")


def test_lbfgsb_quadratic_interior_minimum():
    lb = _lb()
    rng = np.random.default_rng(1)
    Q = rng.standard_normal((8, 8))
    A = Q @ Q.T + 8 * np.eye(8)
    xs = rng.uniform(0.2, 0.8, 8)
    b = A @ xs
    x, f, info = lb(lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b), np.zeros(8), lo=0.0, hi=1.0,
                    eps=1e-10, eps_rel=0.0, past=0)
    # the value stops resolving the step near 1e-8 (as scipy's Fortran code)
    np.testing.assert_allclose(x, xs, atol=1e-7)
    x, f, info = lb(lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b), np.zeros(8), lo=0.0, hi=1.0)
    assert info["reason"].startswith("converged")
    np.testing.assert_allclose(x, xs, atol=1e-4)


@pytest.mark.parametrize("seed", range(6))
def test_lbfgsb_random_box_qp_matches_active_set_solution(seed):
    lb = _lb()
    rng = np.random.default_rng(seed)
    n = 30
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T / n + 0.1 * np.eye(n)
    b = rng.standard_normal(n) * 2.0
    lo, hi = np.full(n, -0.5), np.full(n, 0.7)
    xs = qp_active_set(A, b, lo, hi)
    assert np.sum((xs <= lo + 1e-12) | (xs >= hi - 1e-12)) >= 3  # bounds really active
    x, f, info = lb(lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b), np.zeros(n), lo=lo, hi=hi,
                    eps=1e-10, eps_rel=0.0, max_ls=40, past=0)
    np.testing.assert_allclose(x, xs, atol=1e-6)
    # KKT at the result: projected gradient zero
    assert info["pgnorm"] <= 1e-6
    assert np.all(x >= lo) and np.all(x <= hi)


def test_lbfgsb_rosenbrock_unbounded_and_with_active_bound():
    lb = _lb()
    x, f, info = lb(rosen, np.array([-1.2, 1.0]), eps=1e-10, eps_rel=0.0, max_ls=40)
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-6)
    assert info["n_eval"] < 80
    # x0 <= 0.5 active: min at (0.5, 0.25)
    x, f, info = lb(rosen, np.array([-1.2, 1.0]), lo=[-2.0, -2.0], hi=[0.5, 2.0], eps=1e-10,
                    eps_rel=0.0, max_ls=40)
    np.testing.assert_allclose(x, [0.5, 0.25], atol=1e-7)


def test_lbfgsb_is_quasi_newton_on_large_problems():
    """Evaluation counts on 100-D Rosenbrock against scipy's Fortran
    L-BFGS-B (the published 1997/2011 code), memory 6.  Unbounded, both take
    ~600; on the box [-1, 0.9] (94 bounds active at the solution) with the
    Fortran code's first step to the Cauchy point both take a few dozen.
    With LBFGSpp's unit-length first step the same box takes ~700: the first
    step does not reach the corner, and the valley is then followed inside
    the box, as in the unbounded case."""
    sp = pytest.importorskip("scipy.optimize")
    lb = _lb()
    n = 100
    x0 = np.full(n, -0.3)
    opts = dict(maxcor=6, gtol=1e-8, ftol=0.0, maxiter=5000, maxfun=20000)
    x, f, info = lb(rosen, x0, eps=1e-8, eps_rel=0.0, max_ls=40, max_iter=5000, past=0)
    r = sp.minimize(rosen, x0, jac=True, method="L-BFGS-B", options=opts)
    np.testing.assert_allclose(x, 1.0, atol=1e-6)
    assert info["n_eval"] <= 1.3 * r.nfev, (info["n_eval"], r.nfev)
    lo, hi = np.full(n, -1.0), np.full(n, 0.9)
    r = sp.minimize(rosen, x0, jac=True, method="L-BFGS-B", bounds=list(zip(lo, hi)), options=opts)
    for first, cap in (("cauchy", 2.0 * r.nfev + 10), ("unit", 1000)):
        x, f, info = lb(rosen, x0, lo=lo, hi=hi, eps=1e-8, eps_rel=0.0, max_ls=40, max_iter=5000,
                        first_step=first)
        assert abs(f - r.fun) <= 1e-9 * max(1.0, abs(r.fun)), first
        # stopped by the objective-change test (delta 1e-10): x to ~1e-5
        np.testing.assert_allclose(x, r.x, atol=3e-5)
        assert info["n_eval"] <= cap, (first, info["n_eval"], r.nfev)


def test_lbfgsb_first_step_is_unit_length_along_the_cauchy_direction():
    """LBFGSpp's first trial: x0 + (xcp - x0) / ||xcp - x0||, with xcp the
    projection of x0 - g0 (the Cauchy point of the identity model when no
    bound is crossed)."""
    lb = _lb()
    pts = []
    c = np.array([3.0, -1.0, 0.5])

    def fg(x):
        pts.append(x.copy())
        return float(np.sum((x - c) ** 2)), 2.0 * (x - c)

    lb(fg, np.zeros(3), lo=-10.0, hi=10.0, max_iter=1)
    d = pts[1] - pts[0]
    assert abs(np.linalg.norm(d) - 1.0) <= 1e-12
    np.testing.assert_allclose(d, c / np.linalg.norm(c), atol=1e-12)


def test_lbfgsb_stop_callback_and_non_finite_region():
    lb = _lb()
    # the objective is undefined beyond x > 2: the line search must step back
    def fg(x):
        if np.any(x > 2.0):
            return np.nan, np.zeros_like(x)
        return float(np.sum((x - 1.5) ** 2)), 2.0 * (x - 1.5)

    x, f, info = lb(fg, np.full(3, -3.0), eps=1e-9, eps_rel=0.0, max_ls=40)
    np.testing.assert_allclose(x, 1.5, atol=1e-7)
    # the restart policy and a single subspace pass reach the same point
    for kw in (dict(restart=True), dict(max_submin=1)):
        x, f, info = lb(fg, np.full(3, -3.0), eps=1e-9, eps_rel=0.0, max_ls=40, **kw)
        np.testing.assert_allclose(x, 1.5, atol=1e-7)
    # stop once the value falls below 1
    x, f, info = lb(lambda x: (float(x @ x), 2 * x), np.full(4, 3.0), stop=lambda x, J: J < 1.0)
    assert info["reason"] == "stopped by the caller" and f < 1.0
    assert info["trace"][-1][1] == f


def test_lbfgsb_errors():
    lb = _lb()
    with pytest.raises(RuntimeError):
        lb(lambda x: (0.0, x), np.zeros(2), lo=1.0, hi=0.0)
    with pytest.raises(RuntimeError):
        lb(lambda x: (0.0, x), np.zeros(2), max_submin=0)

    def boom(x):
        raise ValueError("objective exploded")

    with pytest.raises(ValueError):
        lb(boom, np.zeros(2))
