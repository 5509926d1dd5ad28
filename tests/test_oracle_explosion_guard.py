"""The oracle's adjoint explosion guard (the reference's watermark rollback,
assim/swSpE_checkpoint.cpp:1035-1117), pinned against an independent
construction: when the guard stops the reverse sweep, the kept gradient and
cost must be exactly those of the sub-problem that starts from the frozen
state at the first kept step -- the function the reference's truncated
replay measures (:1355-1400)."""
import copy

import numpy as np
import pytest


def _problem(sw, oracle):
    from sw2d_amd import calibration
    cal = calibration.Calibration(px=16, window=300.0)
    E = oracle.OracleE(cal.case)
    K = cal.K(cal.n_default)
    rec = E.record(list(cal.obs_t), K1=K, K2=K)
    rng = np.random.default_rng(0)
    obs = rng.uniform(0.0, 1.0, (len(cal.obs_t), len(cal.gauge_cells)))
    return cal, E, K, rec, obs


def _subproblem(cal, E, K, rec, obs, n0, oracle):
    """Exact gradient of the cost from step n0 on, the state at n0 frozen."""
    head = {"dt": rec["dt"][:n0], "bc": rec["bc"][:n0], "trig": rec["trig"][:n0]}
    _, (H, QX, QY) = E.replay(head, K1=K, K2=K)
    sub = copy.copy(cal.case)
    sub.a = dict(cal.case.a)
    sub.a["h0"], sub.a["qx0"], sub.a["qy0"] = H, QX, QY
    tail = {"dt": rec["dt"][n0:], "bc": rec["bc"][n0:], "trig": rec["trig"][n0:]}
    return oracle.OracleE(sub).gradient(tail, K1=K, K2=K, gauges=cal.gauge_cells, obs=obs,
                                        n_obs=len(cal.obs_t))


@pytest.mark.parametrize("ratio", [1e-9, 1.0, 2.0])
def test_oracle_guard_equals_frozen_subproblem(sw, oracle, ratio):
    cal, E, K, rec, obs = _problem(sw, oracle)
    n = len(rec["dt"])
    S = max(2, n // 12)
    kw = dict(K1=K, K2=K, gauges=cal.gauge_cells, obs=obs, n_obs=len(cal.obs_t))
    G = E.gradient(rec, explosion=(ratio, S), **kw)
    n0 = G["first_kept_step"]
    if ratio < 1e-6:  # any non-zero adjoint trips it: stop after the last segment
        assert n0 == ((n - 1) // S) * S
    if n0 == 0:
        full = E.gradient(rec, **kw)
        for key in ("phi", "K1", "K2", "bc"):
            np.testing.assert_array_equal(G[key], full[key])
        return
    assert n0 % S == 0 and 0 < n0 < n
    Gs = _subproblem(cal, E, K, rec, obs, n0, oracle)
    assert G["J"] == Gs["J"]
    for key in ("phi", "K1", "K2"):
        np.testing.assert_array_equal(G[key], Gs[key])
    np.testing.assert_array_equal(G["bc"][n0:], Gs["bc"])
    assert not np.any(G["bc"][:n0])


def test_oracle_guard_off_is_the_exact_sweep(sw, oracle):
    cal, E, K, rec, obs = _problem(sw, oracle)
    kw = dict(K1=K, K2=K, gauges=cal.gauge_cells, obs=obs, n_obs=len(cal.obs_t))
    full = E.gradient(rec, **kw)
    off = E.gradient(rec, explosion=(0.0, 7), **kw)
    high = E.gradient(rec, explosion=(1e300, 7), **kw)
    assert full["first_kept_step"] == off["first_kept_step"] == high["first_kept_step"] == 0
    for key in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(full[key], off[key])
        np.testing.assert_array_equal(full[key], high[key])
    assert full["J"] == off["J"] == high["J"]
