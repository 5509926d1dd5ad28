"""The 68-control calibration workload (calibration.py) on the GPU: the
gradient with respect to the Manning n of each zone and the hydrograph knots,
mapped from the per-cell / per-step adjoint, against central finite
differences of the frozen-schedule cost."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(sw):
    from sw2d_amd import calibration
    cal = calibration.Calibration(px=40, window=300.0)
    s = sw.SpSolver(cal.case)
    obs = cal.observe(s, cal.n_true, cal.Q_true)
    return cal, s, obs


def test_gpu_zone_and_knot_gradient_vs_fd(sw, small):
    from sw2d_amd import assim
    cal, s, obs = small
    print("[gpu] calibration gradient", cal.case.n_cell, "cells", flush=True)
    n0, Q0 = cal.n_default, cal.Q_default
    cal.apply_hydrograph(s, Q0)
    A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs)
    K = cal.K(n0)
    G = A.gradient(K1=K, K2=K)
    rec = A.get_record()
    gn, gQ = cal.grad_controls(G, n0, rec["t"][:-1])
    assert G["J"] > 0 and np.any(gn) and np.any(gQ)
    for z in np.argsort(-np.abs(gn))[:4]:
        e = 1e-6 * n0[z]
        Kp, Km = cal.K(n0 + e * (np.arange(8) == z)), cal.K(n0 - e * (np.arange(8) == z))
        fd = (A.cost(K1=Kp, K2=Kp) - A.cost(K1=Km, K2=Km)) / (2 * e)
        assert abs(fd - gn[z]) <= 1e-5 * abs(gn[z]) + 1e-12, (z, fd, gn[z])
    W = assim.interp_weights(cal.t_knots, rec["t"][:-1])
    bc0 = rec["bc"].copy()
    for k in np.argsort(-np.abs(gQ))[:4]:
        e = 1e-3
        d = np.zeros_like(bc0)
        d[:, cal.series] = W[:, k] / cal.west_len
        A.set_bc_seq(bc0 + e * d)
        Jp = A.cost(K1=K, K2=K)
        A.set_bc_seq(bc0 - e * d)
        Jm = A.cost(K1=K, K2=K)
        fd = (Jp - Jm) / (2 * e)
        assert abs(fd - gQ[k]) <= 1e-5 * abs(gQ[k]) + 1e-12, (k, fd, gQ[k])
    A.set_bc_seq(bc0)


def test_gpu_fourdvar_recovers_twin_controls(sw):
    """Twin experiment on a small wet-start mesh: L-BFGS from the default
    controls drives the gauge misfit down by 3 orders of magnitude and
    recovers the zone roughness."""
    from sw2d_amd import calibration
    cal = calibration.Calibration(px=40, window=600.0, level_quantile=0.9)
    s = sw.SpSolver(cal.case)
    print("[gpu] 4D-Var twin", cal.case.n_cell, "cells", flush=True)
    obs = cal.observe(s, cal.n_true, cal.Q_true)
    n, Q, trace = cal.fourdvar(s, obs, cal.n_default, cal.Q_default, max_iter=40)
    J0 = trace[0][1]
    Jbest = min(J for _, J, kind in trace if kind == "grad")
    assert Jbest <= 1e-3 * J0, (J0, Jbest)
    # zones with gauges in them are identified; all stay in the bounds
    assert np.all((n >= 0.01) & (n <= 0.1))
    assert np.median(np.abs(n - cal.n_true) / cal.n_true) < 0.05
