"""The oracle's MUSCL-EVR reverse sweep (orc_e_gradient on a muscl_evr case:
recon_adj, lr_bnd_muscl_adj and e_slopes_adj in oracle/sp_oracle_e.c)
against central differences of the oracle's own frozen replay, which the
e_*_muscl_* fixtures pin bit-exactly to the reference's _E forward.

Central differences are taken over windows where they converge: the
second-order scheme switches its limiter, eigenvalue clips and wet/dry
offsets often, and over long windows (the 991-step e_scale_muscl_strickler
run) those kinks make them wander by 1e-5..1e-1 while the Godunov run on the
same mesh stays at 1e-8; its first 300 steps are used here.  Where a wet/dry
front moves the cost is not even continuous (the negative-depth clamp zeroes
the discharge, HLLC changes regime): the dam break is checked over its first
40 steps, where differences at 1e-7 and 1e-8 agree.

Also pinned: the planted mistakes (the slopes' adjoint dropped; the reverse
reconstructing with the dry run's dt) fail the same checks, and the
checkpoint restarts of the reverse sweep (orc_e_set_checkpoints): the dry
run's states on a given schedule are its recorded ones, Godunov gradients do
not depend on the stride, MUSCL-EVR ones do, as the reference's segment
replay from dry-run checkpoints makes them (swSpE_checkpoint.cpp:985-1030).
"""
import numpy as np
import pytest

from conftest import regen_case


def _setup(sw, oracle, tmp_path_factory, name, obs_t, n_steps=None):
    d = str(tmp_path_factory.mktemp(name))
    regen_case(sw, name, d)
    case = sw.Case(d)
    E = oracle.OracleE(case)
    rec = E.record(obs_t)
    if n_steps is not None:  # a shorter window of the recorded schedule
        rec = {k: v[:n_steps].copy() for k, v in rec.items()}
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][::7][:6]
    obs = np.array([[h0[q] * (1 + 0.1 * (k + 1)) + 0.01 for q in cells] for k in range(len(obs_t))])
    kw = dict(gauges=cells, obs=obs, normalize=True, n_obs=len(obs_t))
    return case, E, rec, kw


@pytest.fixture(scope="module")
def mixed(sw, oracle, tmp_path_factory):
    return _setup(sw, oracle, tmp_path_factory, "e_mixed_muscl_strickler", [5.0, 10.0, 15.0, 20.0])


def _dir_fd(E, rec, kw, key, x0, v, e):
    J = lambda x: E.replay(rec, **{key: x}, **kw)[0]
    return (J(x0 + e * v) - J(x0 - e * v)) / (2 * e)


def _controls(case):
    return {"phi": case.a["phi"], "K1": case.a["fc1"], "K2": case.a["fc2"]}


@pytest.mark.parametrize("key", ["phi", "K1", "K2"])
def test_muscl_cell_controls_vs_central_fd(mixed, key):
    case, E, rec, kw = mixed
    G = E.gradient(rec, **kw)
    x0 = _controls(case)[key]
    rng = np.random.default_rng({"phi": 1, "K1": 2, "K2": 3}[key])
    v = np.where(case.a["phi"] > 0.05, rng.standard_normal(case.n_cell), 0.0) * np.maximum(np.abs(x0), 1e-3)
    fd = _dir_fd(E, rec, kw, key, x0, v, 1e-7)
    adj = float(np.dot(G[key], v))
    assert abs(fd - adj) <= 1e-6 * abs(adj) + 4e-16 * abs(G["J"]) / 1e-7, (key, fd, adj)


@pytest.mark.parametrize("series", [1, 2, 3, 4])  # q inflow, fr outlet, h level, z level
def test_muscl_boundary_values_vs_central_fd(mixed, series):
    case, E, rec, kw = mixed
    G = E.gradient(rec, **kw)
    bc0 = rec["bc"].copy()
    d = np.zeros_like(bc0)
    d[:, series] = np.linspace(0.0, 1.0, len(bc0))
    e = 1e-6
    fd = (E.replay(rec, bc=bc0 + e * d, **kw)[0] - E.replay(rec, bc=bc0 - e * d, **kw)[0]) / (2 * e)
    adj = float(np.sum(G["bc"] * d))
    assert abs(fd - adj) <= 1e-6 * abs(adj), (series, fd, adj)


@pytest.mark.parametrize("name,obs_t,n_steps", [
    ("e_dambreak_muscl", [2.5, 5.0, 7.5, 10.0], 40),
    ("e_drybreak_muscl", [2.5, 5.0, 7.5, 10.0], None),
    ("e_scale_muscl_strickler", [30.0, 60.0, 90.0, 120.0], 300),
])
def test_muscl_porosity_direction_vs_central_fd(sw, oracle, tmp_path_factory, name, obs_t, n_steps):
    case, E, rec, kw = _setup(sw, oracle, tmp_path_factory, name, obs_t, n_steps)
    if n_steps is not None:  # the truncated window ends with an observation
        rec["trig"][-1] = max(rec["trig"][-1], 0)
    G = E.gradient(rec, **kw)
    phi = case.a["phi"]
    v = np.where(phi > 0.05, np.random.default_rng(5).standard_normal(case.n_cell), 0.0)
    fd = _dir_fd(E, rec, kw, "phi", phi, v, 1e-7)
    adj = float(np.dot(G["phi"], v))
    assert abs(fd - adj) <= 1e-6 * abs(adj), (name, fd, adj)


@pytest.mark.parametrize("plant", [1, 2])
def test_muscl_planted_mistakes_fail_the_check(mixed, oracle, plant):
    """1: slopes' adjoint dropped; 2: the reverse reconstructs with the
    previous step's dt.  Each moves the gradient far outside the bar."""
    case, E, rec, kw = mixed
    x0 = case.a["phi"]
    v = np.where(x0 > 0.05, np.random.default_rng(1).standard_normal(case.n_cell), 0.0)
    fd = _dir_fd(E, rec, kw, "phi", x0, v, 1e-7)
    L = oracle._e_decl()
    L.orc_e_set_plant(plant)
    try:
        G = E.gradient(rec, **kw)
    finally:
        L.orc_e_set_plant(0)
    adj = float(np.dot(G["phi"], v))
    assert abs(fd - adj) > 1e-4 * abs(fd), (plant, fd, adj)


@pytest.mark.parametrize("name", ["e_mixed_godunov_strickler", "e_mixed_muscl_strickler"])
def test_checkpoint_restarts(sw, oracle, tmp_path_factory, name):
    obs_t = [5.0, 10.0, 15.0, 20.0]
    case, E, rec0, kw = _setup(sw, oracle, tmp_path_factory, name, obs_t)
    rec = E.record(obs_t, states=True)
    S = E.dry_states(rec)
    np.testing.assert_array_equal(S, rec["states"])
    G0 = E.gradient(rec, **kw)
    Gall = E.gradient(rec, checkpoints=(10 ** 6, S), **kw)  # one segment: no restart
    for k in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(G0[k], Gall[k])
    G5 = E.gradient(rec, checkpoints=(5, S), **kw)
    d = np.max(np.abs(G5["phi"] - G0["phi"])) / np.max(np.abs(G0["phi"]))
    if "godunov" in name:  # dry run and replay are one trajectory
        assert d == 0.0 and G5["J"] == G0["J"]
    else:  # restarts from the dry run's states (previous-step dt) move it
        assert d > 1e-6 and G5["J"] != G0["J"]
