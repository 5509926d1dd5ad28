"""The porosity change of variable in the GPU 4D-Var (sw2dg_assim_set_phi_transform):
with a transform the control is m, phi = correct(m) is applied before the
dry run and dJ/dm = dJ/dphi dphi/dm after it -- equal to the untransformed
gradient at phi = correct(m) times the derivative, and the normalisation's
bounds stay frozen for the costs that follow a dry run."""
import numpy as np
import pytest

from test_gpu_adjoint import OBS_T, _gauges, _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,lam", [("tanh", 2.0), ("normalize", 1.0), ("clamp", 1.0)])
def test_gpu_phi_transform_is_the_chain_rule(sw, tmp_path, mode, lam):
    from sw2d_amd import assim
    g, case, s = _setup(sw, str(tmp_path))
    cells, obs = _gauges(case)
    print(f"[gpu] phi transform {mode}", flush=True)
    rng = np.random.default_rng(0)
    phi0 = np.clip(case.a["phi"], 0.05, 0.95)
    if mode == "tanh":
        m = np.log(phi0 / (1.0 - phi0)) / lam
    else:
        m = phi0 + 0.05 * rng.standard_normal(len(phi0))
    m2 = m + 0.01 * rng.standard_normal(len(m))
    A = assim.Assim(s, OBS_T, gauge_cells=cells, gauge_h=obs, phi_transform=(mode, lam))
    G1 = A.gradient(phi=m)
    J1b = A.cost(phi=m2)  # frozen bounds of the dry run at m
    lo, hi = float(m.min()), float(m.max())
    phi, dphi = assim.phi_transform_eval(mode, m, lam=lam, lo=lo, hi=hi)
    phi2, _ = assim.phi_transform_eval(mode, m2, lam=lam, lo=lo, hi=hi)
    B = assim.Assim(s, OBS_T, gauge_cells=cells, gauge_h=obs)
    G0 = B.gradient(phi=phi)
    J0b = B.cost(phi=phi2)
    assert G1["J"] == G0["J"]
    assert J1b == J0b
    np.testing.assert_array_equal(G1["phi"], G0["phi"] * dphi)
    np.testing.assert_array_equal(G1["K1"], G0["K1"])
    s.close()


def test_gpu_porosity_problem_descends(sw, tmp_path):
    """The reference's porosity problem (ridge + spatial penalties, logistic
    change of variable) over the GPU gradient: grad() is consistent with
    central differences of func() along its own direction, and the projected
    L-BFGS lowers f."""
    from sw2d_amd import assim
    from sw2d_amd.porosity import PorosityProblem, minimize
    g, case, s = _setup(sw, str(tmp_path))
    cells, obs = _gauges(case)
    print("[gpu] porosity problem", case.n_cell, "cells", flush=True)
    A = assim.Assim(s, OBS_T, gauge_cells=cells, gauge_h=obs, phi_transform=("tanh", 2.0))
    phi_b = np.clip(case.a["phi"], 0.05, 0.95)
    m_b = np.log(phi_b / (1.0 - phi_b)) / 2.0
    P = PorosityProblem(A, m_b, alpha=1e-3, alpha_spatial=1e-3,
                        edges=(case.a["int_L"], case.a["int_R"]), phi_transform=("tanh", 2.0))
    th0 = P.theta0()
    J0, g0 = P.grad(th0)
    d = -g0 / np.linalg.norm(g0)
    e = 1e-6
    fd = (P.func(th0 + e * d) - P.func(th0 - e * d)) / (2 * e)
    assert fd == pytest.approx(np.dot(g0, d), rel=1e-4)
    x, J, tr = minimize(P.grad, P.func, th0, lo=-5.0, hi=5.0, max_iter=5)
    assert P.func(x) < P.func(th0)
    s.close()


@pytest.mark.parametrize("method", ["gd", "adam", "bfgs", "lbfgsb"])
def test_gpu_run_minimization_methods_descend(sw, tmp_path, method):
    """Each of the reference's minimiser methods (run_minimization) lowers
    the porosity problem's f on the GPU gradient in a few iterations; under
    the clamp change of variable the projected ones keep m in [0, 1]."""
    from sw2d_amd import assim
    from sw2d_amd.porosity import PorosityProblem, run_minimization
    g, case, s = _setup(sw, str(tmp_path))
    cells, obs = _gauges(case)
    print(f"[gpu] run_minimization {method}", flush=True)
    A = assim.Assim(s, OBS_T, gauge_cells=cells, gauge_h=obs, phi_transform=("clamp", 1.0))
    phi_b = np.clip(case.a["phi"], 0.05, 0.95)
    P = PorosityProblem(A, phi_b, alpha=1e-3, phi_transform=("clamp", 1.0), zonation=True)
    th0 = P.theta0()
    f0 = P.grad(th0)[0]  # a dry run at th0 (func replays the last dry run's schedule)
    kw = {"step": 1e-3} if method in ("gd", "adam") else {}
    th, J, tr = run_minimization(P, method, max_iter=4, **kw)
    assert P.grad(th)[0] < f0
    if method != "bfgs":
        assert np.all(th >= 0.0) and np.all(th <= 1.0)
    s.close()


def test_gpu_qsva_continuation_descends_every_window(sw, tmp_path):
    """QSVA (run_qsva) on the GPU porosity problem: windows [0, 10], [0, 20]
    (delta_tau 10 s over the observations at 5, 10, 15, 20 s), each window's
    minimisation starts from the previous optimum and lowers its own f, and
    the final porosity stays in [0, 1] under the clamp change of variable."""
    from sw2d_amd import assim
    from sw2d_amd.porosity import PorosityProblem, run_qsva
    g, case, s = _setup(sw, str(tmp_path))
    cells, obs = _gauges(case)
    t_obs = np.asarray(OBS_T)
    print("[gpu] qsva", case.n_cell, "cells", flush=True)
    starts = []

    def make(t_end, phi):
        k = int(np.sum(t_obs <= t_end + 1e-9))
        A = assim.Assim(s, t_obs[:k], gauge_cells=cells, gauge_h=obs[:k],
                        phi_transform=("clamp", 1.0))
        P = PorosityProblem(A, phi, alpha=1e-3, phi_transform=("clamp", 1.0), zonation=True)
        starts.append((P, P.grad(P.theta0())[0]))
        return P

    phi0 = np.clip(case.a["phi"], 0.05, 0.95)
    phi, steps = run_qsva(make, 0.0, 20.0, 10.0, phi0, method="lbfgsb", max_iter=3)
    assert [st["t_end"] for st in steps] == [10.0, 20.0]
    for (P, f0), st in zip(starts, steps):
        assert st["J"] < f0
    np.testing.assert_array_equal(starts[1][0].phi_b, steps[0]["phi"])   # warm start
    assert np.all((phi >= 0.0) & (phi <= 1.0))
    s.close()
