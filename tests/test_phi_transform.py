"""The porosity change of variable (sw2dg_phi_transform_eval, host code of the
library: phi_correct / dphi_dm of assim/swSpE.h:59-107): known values of
each mode, the derivative against central differences of the map, and the
normalisation's degenerate guard (swSpE_core.cpp:229-266)."""
import numpy as np
import pytest


def _ev(sw, mode, m, **kw):
    from sw2d_amd import assim
    return assim.phi_transform_eval(mode, m, **kw)


def test_known_values(sw):
    m = np.array([-0.5, 0.0, 0.3, 1.05, 1.2])
    phi, d = _ev(sw, "clamp", m)
    np.testing.assert_array_equal(phi, [0.0, 0.0, 0.3, 1.05, 1.05])
    np.testing.assert_array_equal(d, [0.0, 1.0, 1.0, 1.0, 0.0])
    phi, d = _ev(sw, "tanh", np.array([0.0]), lam=4.0)
    assert phi[0] == 0.5 and d[0] == 1.0       # lambda s (1 - s) at s = 1/2
    phi, d = _ev(sw, "normalize", np.array([2.0, 3.0, 6.0]), lo=2.0, hi=6.0)
    np.testing.assert_allclose(phi, [0.0, 0.25, 1.0], rtol=0, atol=1e-16)
    np.testing.assert_allclose(d, 0.25, rtol=1e-15)
    # uniform map: normalisation off (phi = m, derivative 1)
    phi, d = _ev(sw, "normalize", np.array([0.7]), lo=0.7, hi=0.7)
    assert phi[0] == 0.7 and d[0] == 1.0
    phi, d = _ev(sw, None, m)
    np.testing.assert_array_equal(phi, m)
    np.testing.assert_array_equal(d, 1.0)


@pytest.mark.parametrize("mode,kw", [("tanh", {"lam": 3.0}), ("normalize", {"lo": -1.0, "hi": 2.5}),
                                     ("clamp", {})])
def test_derivative_matches_differences(sw, mode, kw):
    rng = np.random.default_rng(0)
    m = rng.uniform(-1.0, 1.0, 64)
    if mode == "clamp":  # away from the kinks at 0 and 1.05
        m = m[(np.abs(m) > 1e-3) & (np.abs(m - 1.05) > 1e-3)]
    e = 1e-6
    _, d = _ev(sw, mode, m, **kw)
    fd = (_ev(sw, mode, m + e, **kw)[0] - _ev(sw, mode, m - e, **kw)[0]) / (2 * e)
    np.testing.assert_allclose(d, fd, rtol=1e-7, atol=1e-9)
