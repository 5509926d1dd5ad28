"""The adjoint explosion guard on the GPU (sw2dg_assim_set_explosion_guard,
the reference's watermark rollback, assim/swSpE_checkpoint.cpp:1035-1117):

- forced (ratio 1e-9, 1, 2): the truncation point, the kept gradient of every
  control kind and the truncated cost equal the oracle's guard on the same
  recorded schedule and segments (itself pinned against the frozen
  sub-problem in test_oracle_explosion_guard.py);
- after a truncation the frozen cost starts at the first kept segment and is
  the J the gradient reported, and central differences of it match the
  kept gradient (what the reference's truncated replay is for, :1355-1400);
- default ratio 1e8 on a case that does not explode: identical to the guard
  off; on the 45k-cell 1200 s case where the exact adjoint explodes, the
  default guard truncates, far below the unguarded adjoint.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(sw):
    from sw2d_amd import calibration
    cal = calibration.Calibration(px=16, window=300.0)
    s = sw.SpSolver(cal.case)
    rng = np.random.default_rng(0)
    obs = rng.uniform(0.0, 1.0, (len(cal.obs_t), len(cal.gauge_cells)))
    return cal, s, obs


@pytest.mark.parametrize("ratio", [1e-9, 1.0, 2.0])
def test_gpu_guard_matches_oracle(sw, oracle, ratio):
    from sw2d_amd import assim
    cal, s, obs = _problem(sw)
    print(f"[gpu] explosion guard ratio {ratio}", flush=True)
    K = cal.K(cal.n_default)
    A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=200,
                    stride=4, explosion_ratio=ratio)
    G = A.gradient(K1=K, K2=K)
    n0, seg = A.truncation()
    rec = A.get_record()
    E = oracle.OracleE(cal.case)
    O = E.gradient(rec, K1=K, K2=K, gauges=cal.gauge_cells, obs=obs, n_obs=len(cal.obs_t),
                   explosion=(ratio, 4))
    assert n0 == O["first_kept_step"] and n0 == 4 * seg
    if ratio < 1e-6:
        assert n0 > 0
    assert abs(G["J"] - O["J"]) <= 1e-12 * abs(O["J"])
    for key in ("phi", "K1", "K2", "bc"):
        scale = np.max(np.abs(O[key]))
        if scale == 0.0:
            assert not np.any(G[key]), key
            continue
        assert np.max(np.abs(G[key] - O[key])) <= 1e-10 * scale, key
    assert not np.any(G["bc"][:n0])
    s.close()


def test_gpu_truncated_cost_is_consistent_with_the_kept_gradient(sw):
    from sw2d_amd import assim
    cal, s, obs = _problem(sw)
    K = cal.K(cal.n_default)
    A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=200,
                    stride=4, explosion_ratio=1.0)
    G = A.gradient(K1=K, K2=K)
    n0, _ = A.truncation()
    assert n0 > 0
    assert A.cost(K1=K, K2=K) == G["J"]
    d = G["K1"] / np.linalg.norm(G["K1"])
    e = 1e-6 * np.max(K)
    fd = (A.cost(K1=K + e * d, K2=K) - A.cost(K1=K - e * d, K2=K)) / (2 * e)
    assert abs(fd - np.dot(G["K1"], d)) <= 1e-5 * abs(np.dot(G["K1"], d)), (fd, np.dot(G["K1"], d))
    s.close()


def test_gpu_default_guard_changes_nothing_without_explosion(sw):
    from sw2d_amd import assim
    cal, s, obs = _problem(sw)
    K = cal.K(cal.n_default)
    out = []
    for ratio in (0.0, 1e8):
        A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=200,
                        stride=4, explosion_ratio=ratio)
        out.append(A.gradient(K1=K, K2=K))
        assert A.truncation() == (0, 0)
    for key in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(out[0][key], out[1][key])
    assert out[0]["J"] == out[1]["J"]
    s.close()


def test_gpu_default_guard_truncates_an_exploding_adjoint(sw):
    from sw2d_amd import assim, calibration
    cal = calibration.Calibration(px=150, window=1200.0, level_quantile=0.9)
    s = sw.SpSolver(cal.case)
    obs = cal.observe(s, cal.n_true, cal.Q_true, max_steps=20000)
    print("[gpu] explosion guard on", cal.case.n_cell, "cells", flush=True)
    cal.apply_hydrograph(s, cal.Q_default)
    K = cal.K(cal.n_default)
    B = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=20000,
                    explosion_ratio=0.0)
    GB = B.gradient(K1=K, K2=K)
    g_exact = cal.grad_controls(GB, cal.n_default, B.get_record()["t"][:-1])[0]
    A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=20000)
    G = A.gradient(K1=K, K2=K)
    n0, seg = A.truncation()
    gn = cal.grad_controls(G, cal.n_default, A.get_record()["t"][:-1])[0]
    print(f"[gpu] truncated at step {n0} of {G['n_steps']}; |dJ/dn| {np.max(np.abs(gn)):.3e} "
          f"(unguarded {np.max(np.abs(g_exact)):.3e})", flush=True)
    assert 0 < n0 < G["n_steps"]
    # the watermark holds what the sweep had gathered before the blow-up was
    # detected (the reference's own caveat: the growth is gradual), far below
    # the unguarded adjoint
    assert np.all(np.isfinite(gn)) and np.max(np.abs(gn)) < 1e-6 * np.max(np.abs(g_exact))
    assert A.cost(K1=K, K2=K) == G["J"]
    s.close()


def test_gpu_frozen_first_segment_is_the_subproblem(sw, oracle):
    """A frozen truncation window (sw2dg_assim_set_first_segment, the
    reference's forced_first_seg_): the gradient and the cost start at the
    segment -- equal to the oracle's exact gradient of the sub-problem from
    the frozen state there -- the guard no longer truncates (ratio 1e-9
    would otherwise), and unfreezing restores the full sweep bit for bit."""
    import copy
    from sw2d_amd import assim
    cal, s, obs = _problem(sw)
    print("[gpu] frozen first segment", flush=True)
    K = cal.K(cal.n_default)
    A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=200,
                    stride=4, explosion_ratio=1e-9)
    A.set_first_segment(5)
    G = A.gradient(K1=K, K2=K)
    n0, seg = A.truncation()
    assert (n0, seg) == (20, 5)
    assert A.cost(K1=K, K2=K) == G["J"]
    rec = A.get_record()
    E = oracle.OracleE(cal.case)
    head = {"dt": rec["dt"][:n0], "bc": rec["bc"][:n0], "trig": rec["trig"][:n0]}
    _, (H, QX, QY) = E.replay(head, K1=K, K2=K)
    sub = copy.copy(cal.case)
    sub.a = dict(cal.case.a)
    sub.a["h0"], sub.a["qx0"], sub.a["qy0"] = H, QX, QY
    tail = {"dt": rec["dt"][n0:], "bc": rec["bc"][n0:], "trig": rec["trig"][n0:]}
    O = oracle.OracleE(sub).gradient(tail, K1=K, K2=K, gauges=cal.gauge_cells, obs=obs,
                                     n_obs=len(cal.obs_t))
    assert abs(G["J"] - O["J"]) <= 1e-12 * abs(O["J"])
    for key in ("phi", "K1", "K2"):
        scale = np.max(np.abs(O[key]))
        assert scale > 0.0
        assert np.max(np.abs(G[key] - O[key])) <= 1e-10 * scale, key
    assert np.max(np.abs(G["bc"][n0:] - O["bc"])) <= 1e-10 * np.max(np.abs(O["bc"]))
    assert not np.any(G["bc"][:n0])
    # unfrozen, guard off: the exact sweep, as a fresh Assim gives it
    A.set_first_segment(None)
    B = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=200,
                    stride=4, explosion_ratio=0.0)
    Gb = B.gradient(K1=K, K2=K)
    C_ = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs, max_steps=200,
                     stride=4, explosion_ratio=0.0)
    C_.set_first_segment(3)
    C_.set_first_segment(-1)
    Gc = C_.gradient(K1=K, K2=K)
    for key in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(Gb[key], Gc[key])
    assert C_.truncation() == (0, 0)
    s.close()
