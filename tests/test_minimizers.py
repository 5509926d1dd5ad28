"""The reference's minimiser methods on PorosityProblem (run_minimization,
assim/swSpE_minimize.cpp:725-1017) as host code over any gradient: projected
GD + Armijo, projected Adam, dense BFGS (the L-BFGS one is in
test_porosity_problem.py).  The reference's minimiser does not compile in
this image (its PSQN / LBFGSpp dependencies are absent), so these are pinned
by what the mathematics fixes: known minimisers (with and without the
[0, 1] projection), Adam's closed form under a constant gradient (which a
dropped bias correction breaks), monotone accepted costs, the stop rules."""
import numpy as np
import pytest


def _quad(c, a):
    c, a = np.asarray(c, float), np.asarray(a, float)
    fg = lambda x: (0.5 * float(np.sum(a * (x - c) ** 2)), a * (x - c))
    return fg, (lambda x: fg(x)[0])


def test_gd_reaches_the_projected_minimiser(sw):
    from sw2d_amd.porosity import minimize_gd
    c = np.array([0.3, 1.7, -0.4, 0.8])
    fg, f = _quad(c, [1.0, 3.0, 0.5, 10.0])
    x, J, trace = minimize_gd(fg, f, np.full(4, 0.5), project=True, step=1e-3, max_iter=200)
    np.testing.assert_allclose(x, np.clip(c, 0.0, 1.0), atol=1e-6)
    grads = [v for (_, v, k) in trace if k == "grad"]
    assert all(b <= a for a, b in zip(grads, grads[1:]))  # accepted costs never rise


def test_gd_unprojected_and_stop_rules(sw):
    from sw2d_amd.porosity import minimize_gd
    fg, f = _quad([2.0, -3.0], [1.0, 1.0])
    x, J, _ = minimize_gd(fg, f, np.zeros(2), project=False, step=1e-6, max_iter=100)
    np.testing.assert_allclose(x, [2.0, -3.0], atol=1e-6)
    # a NaN gradient ends the descent at once, on the first (best) iterate
    x, J, tr = minimize_gd(lambda x: (1.0, np.full(2, np.nan)), lambda x: 1.0, np.ones(2))
    assert J == 1.0 and len(tr) == 1
    # a step that never gives a finite decrease: no move
    x, J, _ = minimize_gd(lambda x: (float(x @ x), 2 * x), lambda x: np.nan, np.ones(2), project=False)
    np.testing.assert_array_equal(x, np.ones(2))


def test_adam_constant_gradient_closed_form(sw):
    """With g constant the bias-corrected moments are g and g^2 at every
    iterate, so each coordinate moves by step g / (|g| + eps) exactly."""
    from sw2d_amd.porosity import minimize_adam
    g = np.array([0.3, -2.0, 5e-3])
    fg = lambda x: (float(g @ x), g.copy())
    step, k = 0.01, 7
    xs = []
    minimize_adam(lambda x: (xs.append(x.copy()), fg(x))[1], np.zeros(3), step=step, max_iter=k,
                  project=False)
    expected = [-i * step * g / (np.abs(g) + 1e-8) for i in range(k)]
    np.testing.assert_allclose(np.array(xs), np.array(expected), rtol=1e-12, atol=1e-15)


def test_adam_projection_best_iterate_and_safe_stop(sw):
    from sw2d_amd.porosity import minimize_adam
    fg, _ = _quad([1.4, 0.2], [1.0, 1.0])
    x, J, _ = minimize_adam(fg, np.array([0.5, 0.5]), step=0.05, max_iter=400, project=True)
    np.testing.assert_allclose(x, [1.0, 0.2], atol=2e-2)
    # a gradient that jumps 100x at the 4th iterate: a safe run stops there
    # (returning the best iterate so far), an unsafe one records it and goes on
    def jumpy():
        calls = []

        def fg(x):
            calls.append(1)
            return 10.0 - len(calls), np.ones(2) * (100.0 if len(calls) == 4 else 1.0)
        return fg, calls

    fg, calls = jumpy()
    x, J, tr = minimize_adam(fg, np.full(2, 0.5), step=0.01, max_iter=10, safe=True)
    assert len(calls) == 4 and J == 6.0
    fg, calls = jumpy()
    x, J, tr = minimize_adam(fg, np.full(2, 0.5), step=0.01, max_iter=10)
    assert len(calls) == 10 and J == 0.0
    assert [k for (_, _, k) in tr].count("explosion") == 1


def test_bfgs_rosenbrock_and_quadratic(sw):
    from sw2d_amd.porosity import minimize_bfgs

    def rosen(x):
        a, b = x
        J = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        return J, np.array([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])

    x, J, tr = minimize_bfgs(rosen, np.array([-1.2, 1.0]), max_iter=200, rel_eps=1e-16)
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-5)
    rng = np.random.default_rng(0)
    M = rng.standard_normal((6, 6))
    A = M @ M.T + 6 * np.eye(6)
    b = rng.standard_normal(6)
    x, J, tr = minimize_bfgs(lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b), np.zeros(6),
                             max_iter=100, rel_eps=1e-16)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-6, atol=1e-8)


def test_run_minimization_dispatch(sw):
    from sw2d_amd.porosity import run_minimization, METHODS
    assert METHODS == ("lbfgsb", "gd", "adam", "bfgs")

    class Quad:  # the PorosityProblem interface the dispatch uses
        mode = None

        def theta0(self):
            return np.full(3, 0.5)

        def grad(self, th):
            return float(np.sum((th - [0.2, 0.9, 1.3]) ** 2)), 2 * (th - [0.2, 0.9, 1.3])

        def func(self, th):
            return self.grad(th)[0]

    for m in METHODS:
        th, J, _ = run_minimization(Quad(), m, max_iter=300, **({"step": 0.02} if m == "adam" else {}))
        want = [0.2, 0.9, 1.3] if m == "bfgs" else [0.2, 0.9, 1.0]  # BFGS is unconstrained
        np.testing.assert_allclose(th, want, atol=3e-2 if m == "adam" else 1e-5)
    with pytest.raises(ValueError):
        run_minimization(Quad(), "nelder-mead")


def test_qsva_window_ends(sw):
    from sw2d_amd.porosity import qsva_window_ends
    assert qsva_window_ends(0.0, 10.0, 3.0) == [3.0, 6.0, 9.0, 10.0]
    assert qsva_window_ends(5.0, 11.0, 2.0) == [7.0, 9.0, 11.0]
    assert qsva_window_ends(0.0, 1.0, 5.0) == [1.0]
    with pytest.raises(ValueError):
        qsva_window_ends(0.0, 1.0, 0.0)


def test_qsva_follows_the_moving_minimum_with_warm_starts(sw):
    """A cost whose minimum moves with the window end (phi* = t_end / 10 on
    every cell) and whose far-away basin traps a cold start on the full
    window: the continuation reaches the full window's absolute minimum from
    warm starts, each window starting at the previous optimum."""
    from sw2d_amd.porosity import run_qsva

    class Prob:
        mode = None

        def __init__(self, t_end, phi):
            self.target = t_end / 10.0
            self.start = np.asarray(phi, dtype=np.float64).copy()
            self.starts = []

        def theta0(self):
            self.starts.append(self.start.copy())
            return self.start.copy()

        def m_of(self, theta):
            return np.asarray(theta, dtype=np.float64)

        def func(self, x):
            d = x - self.target
            # a double well: a secondary minimum near 1 - target beyond a ridge
            return float(np.sum(d ** 2 * (x - (1.0 - self.target)) ** 2 + 0.01 * d ** 2))

        def grad(self, x):
            d = x - self.target
            e = x - (1.0 - self.target)
            g = 2 * d * e ** 2 + 2 * d ** 2 * e + 0.02 * d
            return self.func(x), g

    made = []

    def make(t_end, phi):
        p = Prob(t_end, phi)
        made.append(p)
        return p

    phi, steps = run_qsva(make, 0.0, 8.0, 1.0, np.full(3, 0.0), method="bfgs", max_iter=200)
    assert [s["t_end"] for s in steps] == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0]
    np.testing.assert_allclose(phi, 0.8, atol=1e-6)
    for k in range(1, len(made)):
        np.testing.assert_allclose(made[k].starts[0], steps[k - 1]["phi"])
    # a cold start from 0 on the full window [0, 8] falls into the other well
    cold = Prob(8.0, np.full(3, 0.0))
    from sw2d_amd.porosity import minimize_bfgs
    x, _, _ = minimize_bfgs(cold.grad, cold.theta0(), max_iter=200)
    assert np.all(np.abs(x - 0.2) < 0.05) and np.all(np.abs(x - 0.8) > 0.5)
