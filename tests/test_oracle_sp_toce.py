"""The reference's functional assimilation case sp_toce (4800 cells, quads and
triangles, head loss, Strickler friction, q/fr boundaries): the oracle's _E
forward at phi = 1 with the 41 reference frames reproduces the cost the
reference reports in the first row of its descent log
(configue_fonctionels/sp_toce/descent_log_reference_-75.8pct.csv)."""
import numpy as np

from conftest import load_sp_toce


def test_oracle_cost_at_phi_one_matches_reference_descent_log(sw, oracle):
    case, g = load_sp_toce(sw)
    E = oracle.OracleE(case)
    ones = np.ones(case.n_cell)
    rec = E.record(g["obs_t"], phi=ones)
    assert int(np.sum(rec["trig"] >= 0)) == len(g["obs_t"])
    J = E.frame_cost(rec, g["frames"], phi=ones, normalize=bool(g["normalize"]))
    # the log prints 11 significant digits
    assert abs(J - float(g["expect_J"])) <= 1e-10 * float(g["expect_J"]), J


def short_horizon(E, g, n=100):
    """Schedule of the first ~n steps of sp_toce at phi = 1 with one frame
    observation (the reference frame of t = 10 s) landing at the end."""
    ones = np.ones(E.nc)
    r0 = E.record([], phi=ones, max_steps=n)
    t_obs = float(np.sum(r0["dt"]))
    rec = E.record([t_obs], phi=ones, max_steps=n + 5)
    k = int(np.nonzero(rec["trig"] == 0)[0][0])
    rec = {key: v[: k + 1] for key, v in rec.items()}
    return rec, t_obs, g["frames"][:1]


def test_oracle_frame_gradient_vs_fd_on_sp_toce(sw, oracle):
    case, g = load_sp_toce(sw)
    E = oracle.OracleE(case)
    rec, t_obs, fr = short_horizon(E, g)
    ones = np.ones(case.n_cell)
    G = E.frame_gradient(rec, fr, phi=ones)
    assert G["J"] == E.frame_cost(rec, fr, phi=ones) and G["J"] > 0
    rng = np.random.default_rng(2)
    d = rng.standard_normal(case.n_cell)
    e = 1e-7
    fd = (E.frame_cost(rec, fr, phi=ones + e * d) - E.frame_cost(rec, fr, phi=ones - e * d)) / (2 * e)
    adj = float(np.dot(G["phi"], d))
    assert abs(fd - adj) <= 1e-6 * abs(adj), (fd, adj)
