"""The oracle's reverse sweep (orc_e_gradient, oracle/sp_oracle_e.c) against
central finite differences of the oracle's own frozen replay, which the e_*
fixtures pin bit-exactly to the reference's _E forward.  Every control kind:
porosity, the two Strickler coefficients per cell, and the boundary value of
each BC type (q inflow, fr outlet, h and z levels)."""
import numpy as np
import pytest

from conftest import regen_case

OBS_T = [5.0, 10.0, 15.0, 20.0]


@pytest.fixture(scope="module")
def setup(sw, oracle, tmp_path_factory):
    d = str(tmp_path_factory.mktemp("adj"))
    regen_case(sw, "e_mixed_godunov_strickler", d)
    case = sw.Case(d)
    E = oracle.OracleE(case)
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][::7][:6]
    obs = np.array([[h0[q] * (1 + 0.1 * (k + 1)) for q in cells] for k in range(len(OBS_T))])
    rec = E.record(OBS_T)
    kw = dict(gauges=cells, obs=obs, normalize=True, n_obs=len(OBS_T))
    G = E.gradient(rec, **kw)
    return case, E, rec, kw, G


def _J(E, rec, kw, **ctl):
    return E.replay(rec, **ctl, **kw)[0]


def test_gradient_reports_the_replay_cost(setup):
    case, E, rec, kw, G = setup
    assert G["J"] == _J(E, rec, kw) and G["J"] > 0


@pytest.mark.parametrize("key", ["phi", "K1", "K2"])
def test_cell_controls_vs_central_fd(setup, key):
    case, E, rec, kw, G = setup
    base = {"phi": case.a["phi"], "K1": case.a["fc1"], "K2": case.a["fc2"]}
    g = G[key]
    idx = np.argsort(-np.abs(np.where(case.a["phi"] > 0.05, g, 0.0)))[:6]
    for i in idx:
        e = 1e-6 * max(abs(base[key][i]), 1e-3)
        xp, xm = base[key].copy(), base[key].copy()
        xp[i] += e
        xm[i] -= e
        fd = (_J(E, rec, kw, **{key: xp}) - _J(E, rec, kw, **{key: xm})) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * abs(g[i]), (key, i, fd, g[i])


def test_random_direction_over_all_porosities(setup):
    """<dJ/dphi, d> for a random direction touching every open cell."""
    case, E, rec, kw, G = setup
    phi = case.a["phi"]
    rng = np.random.default_rng(5)
    d = np.where(phi > 0.05, rng.standard_normal(case.n_cell), 0.0)
    e = 1e-7
    fd = (_J(E, rec, kw, phi=phi + e * d) - _J(E, rec, kw, phi=phi - e * d)) / (2 * e)
    adj = float(np.dot(G["phi"], d))
    assert abs(fd - adj) <= 1e-6 * abs(adj), (fd, adj)


@pytest.mark.parametrize("series", [1, 2, 3, 4])  # q inflow, fr outlet, h level, z level
def test_boundary_values_vs_central_fd(setup, series):
    case, E, rec, kw, G = setup
    bc0 = rec["bc"].copy()
    d = np.zeros_like(bc0)
    d[:, series] = np.linspace(0.0, 1.0, len(bc0))
    e = 1e-6
    fd = (_J(E, rec, kw, bc=bc0 + e * d) - _J(E, rec, kw, bc=bc0 - e * d)) / (2 * e)
    adj = float(np.sum(G["bc"] * d))
    assert abs(fd - adj) <= 1e-6 * abs(adj), (series, fd, adj)


def test_fd_check_detects_a_wrong_gradient(setup):
    """The check above fails for a gradient off by a plausible mistake (a
    dropped factor 2 in the cost seed, i.e. half the gradient)."""
    case, E, rec, kw, G = setup
    phi = case.a["phi"]
    rng = np.random.default_rng(5)
    d = np.where(phi > 0.05, rng.standard_normal(case.n_cell), 0.0)
    e = 1e-7
    fd = (_J(E, rec, kw, phi=phi + e * d) - _J(E, rec, kw, phi=phi - e * d)) / (2 * e)
    wrong = 0.5 * float(np.dot(G["phi"], d))
    assert abs(fd - wrong) > 1e-6 * abs(wrong)
