"""The oracle's thin-film cut of the reverse sweep (test infrastructure for
sw2dg_assim_set_adjoint_cut): a cut of 0 is the exact reverse sweep and a
cut above every depth drops every adjoint (zero gradient, cost unchanged).
The cut in between is pinned by the GPU parity test against this oracle and
by the GPU's agreement with finite differences where the exact adjoint
explodes (tests/test_gpu_adjoint_cut.py)."""
import numpy as np


def _small(sw, oracle):
    from sw2d_amd import calibration
    cal = calibration.Calibration(px=16, window=120.0)
    E = oracle.OracleE(cal.case)
    K = cal.K(cal.n_default)
    rec = E.record(list(cal.obs_t), K1=K, K2=K)
    obs = np.zeros((len(cal.obs_t), len(cal.gauge_cells)))
    return cal, E, K, rec, obs


def test_oracle_cut_zero_is_exact_and_huge_cut_drops_everything(sw, oracle):
    cal, E, K, rec, obs = _small(sw, oracle)
    kw = dict(K1=K, K2=K, gauges=cal.gauge_cells, obs=obs, n_obs=len(cal.obs_t))
    G = E.gradient(rec, **kw)
    G0 = E.gradient(rec, adj_cut=0.0, **kw)
    Gbig = E.gradient(rec, adj_cut=1e300, **kw)
    assert G["J"] == G0["J"] == Gbig["J"] > 0.0
    for key in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(G[key], G0[key])
        assert not np.any(Gbig[key]), key
    assert np.any(G["K1"]) and np.any(G["bc"])
    # the setter does not leak into later calls
    np.testing.assert_array_equal(E.gradient(rec, **kw)["K1"], G["K1"])

