"""Host logic of the learned warm start (no GPU): the encodings between gauge
series / controls and the network's inputs / outputs, and the calibration's
refusal of a run that stopped before the last observation."""
import numpy as np
import pytest


def test_targets_and_controls_are_inverse(sw):
    from sw2d_amd.warmstart import WarmStart
    rng = np.random.default_rng(0)
    Qd = rng.uniform(5.0, 300.0, 60)
    ws = WarmStart(Qd)
    n = rng.uniform(0.015, 0.06, (5, 8))
    Q = Qd * (1.0 + rng.uniform(-0.3, 0.3, (5, 60)))
    y = ws.targets(n, Q)
    assert y.shape == (5, 68)
    np.testing.assert_allclose(y[:, :8], np.log(n / 0.03), rtol=1e-15)
    n2, Q2 = ws.controls(y)
    np.testing.assert_allclose(n2, n, rtol=1e-14)
    np.testing.assert_allclose(Q2, Q, rtol=1e-14)


def test_features_are_gauge_major_depth_changes(sw):
    from sw2d_amd.warmstart import WarmStart
    n_obs, n_g = 40, 20
    g = np.arange(3 * n_obs * n_g, dtype=np.float64).reshape(3, n_obs, n_g)
    h0 = np.arange(3 * n_g, dtype=np.float64).reshape(3, n_g) * 0.5
    x = WarmStart.features(g, h0)
    assert x.shape == (3, n_obs * n_g)
    for r, k, o in ((0, 0, 0), (1, 7, 13), (2, 19, 39)):
        assert x[r, k * n_obs + o] == g[r, o, k] - h0[r, k]
    # one run (no batch axis) encodes like a batch of one
    np.testing.assert_array_equal(WarmStart.features(g[1], h0[1]), x[1])


def test_calibration_refuses_a_truncated_window(sw):
    from sw2d_amd.calibration import Calibration
    cal = Calibration(px=8, window=300.0)
    cal._check_window({"t": np.array([0.0, 300.0])}, 10)
    with pytest.raises(RuntimeError, match="max_steps"):
        cal._check_window({"t": np.array([0.0, 250.0])}, 10)
