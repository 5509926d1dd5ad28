"""4D-Var layer on the GPU: the differentiable (_E) forward, its cost and the
hand-written adjoint, all through the C-ABI (sw2dg_assim_*).

- dry run vs the reference's own _E physics (e_* fixtures made by
  oracle/_ref/ref_e_driver from assim/swSpE_*.cpp): observation triggers
  identical, dt and final state within 1e-12 relative;
- cost on the frozen schedule vs the CPU oracle's replay (1e-12 relative);
- gradients vs central finite differences of the GPU cost (porosity,
  Strickler K1/K2 per cell, boundary values per step), with tiles, window and
  frame observations;
- the checkpoint stride changes nothing (bit-identical gradient), and two
  runs are bit-identical.
"""
import numpy as np
import pytest

from conftest import e_golden_names, regen_case, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12
OBS_T = [5.0, 10.0, 15.0, 20.0]


def _setup(sw, folder, name="e_mixed_godunov_strickler"):
    g = regen_case(sw, name, folder)
    print(f"[gpu] assim case {name}", flush=True)
    case = sw.Case(folder)
    return g, case, sw.SpSolver(case)


def _gauges(case, n=6):
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][:: max(1, int(np.sum(h0 > 0.1)) // n)][:n]
    obs = np.array([[h0[c] * (1.0 + 0.1 * (k + 1)) for c in cells] for k in range(len(OBS_T))])
    return cells.astype(np.int64), obs


def _assim(sw, s, case, **kw):
    from sw2d_amd import assim
    cells, obs = _gauges(case)
    args = dict(gauge_cells=cells, gauge_h=obs, normalize=True)
    args.update(kw)
    return assim.Assim(s, OBS_T, **args), cells, obs


# Cases whose path has a transcendental function that the second-order
# scheme's limiter would amplify if it differed by an ulp (the inflow's
# exp(log(q / phi / g) / 3), the friction's cbrt): with the kernels' glibc
# exp / cbrt and correctly rounded log (csrc/sw_math.h) their dry runs are bit
# for bit the reference's (the few log results that differ from glibc's are
# absorbed by the following exp; tests/test_oracle_math_audit.py)
CHAOTIC = {"e_scale_muscl_strickler", "e_scale_muscl_nofriction", "e_mixed_muscl_strickler"}


@pytest.mark.parametrize("name", e_golden_names())
def test_gpu_dry_run_matches_reference_e(sw, tmp_path, name):
    from sw2d_amd import assim
    g, case, s = _setup(sw, str(tmp_path), name)
    A = assim.Assim(s, list(g["obs_t"]), max_steps=5000)
    n = A.record()
    assert n == len(g["dt"])
    r = A.get_record()
    np.testing.assert_array_equal(r["trig"], g["trig"])
    # the whole run bit for bit the reference's (dt of every step, final state)
    np.testing.assert_array_equal(r["dt"], g["dt"])
    h, qx, qy = A.final_state()
    for got, ref in ((h, g["h"]), (qx, g["qx"]), (qy, g["qy"])):
        np.testing.assert_array_equal(got, ref)
    s.close()


def test_gpu_cost_matches_oracle_replay(sw, oracle, tmp_path):
    g, case, s = _setup(sw, str(tmp_path))
    A, cells, obs = _assim(sw, s, case)
    A.record()
    J = A.cost()
    E = oracle.OracleE(case)
    rec = E.record(OBS_T)
    Jo, _ = E.replay(rec, gauges=cells, obs=obs, normalize=True, n_obs=len(OBS_T))
    assert abs(J - Jo) <= TOL * abs(Jo), (J, Jo)
    # the gradient call reports the same cost (same dry run, same schedule)
    G = A.gradient()
    assert abs(G["J"] - Jo) <= TOL * abs(Jo)
    # perturbed controls on the frozen schedule: still the oracle's cost
    phi = case.a["phi"] * 0.97
    K1 = case.a["fc1"] * 1.1
    J2 = A.cost(phi=phi, K1=K1)
    Jo2, _ = E.replay(rec, phi=phi, K1=K1, gauges=cells, obs=obs, normalize=True, n_obs=len(OBS_T))
    assert abs(J2 - Jo2) <= TOL * abs(Jo2), (J2, Jo2)
    s.close()


def _fd(f, x, idx, rel=1e-6):
    out = []
    for i in idx:
        e = rel * max(abs(x[i]), 1e-3)
        xp, xm = x.copy(), x.copy()
        xp[i] += e
        xm[i] -= e
        out.append((f(xp) - f(xm)) / (2 * e))
    return np.array(out)


def _check(fd, adj, scale, rtol=2e-5):
    err = np.abs(fd - adj)
    assert np.all(err <= rtol * np.abs(adj) + 1e-9 * scale), np.c_[fd, adj, err]


def test_gpu_gradient_porosity_and_friction_vs_fd(sw, tmp_path):
    g, case, s = _setup(sw, str(tmp_path))
    A, _, _ = _assim(sw, s, case)
    phi0, K10, K20 = (case.a[k].copy() for k in ("phi", "fc1", "fc2"))
    G = A.gradient()
    assert G["n_steps"] == len(g["dt"])
    for key, x0 in (("phi", phi0), ("K1", K10), ("K2", K20)):
        gr = G[key]
        assert np.all(np.isfinite(gr))
        ok = phi0 > 0.05  # open cells (closed ones sit on the phi_min branch)
        idx = np.argsort(-np.abs(np.where(ok, gr, 0.0)))[:6]
        assert np.abs(gr[idx[0]]) > 0

        def J(x, key=key):
            c = {"phi": phi0, "K1": K10, "K2": K20}
            c[key] = x
            return A.cost(phi=c["phi"], K1=c["K1"], K2=c["K2"])

        _check(_fd(J, x0, idx), gr[idx], np.max(np.abs(gr)))
    s.close()


def test_gpu_gradient_boundary_values_directional_fd(sw, tmp_path):
    """dJ/d(boundary value per step): directional derivative along a
    hydrograph-knot direction on the frozen schedule."""
    from sw2d_amd import assim
    g, case, s = _setup(sw, str(tmp_path))
    A, _, _ = _assim(sw, s, case)
    G = A.gradient()
    r = A.get_record()
    bc0 = r["bc"].copy()
    t_steps = r["t"][:-1]
    W = assim.interp_weights(case.a["series_t"], t_steps)
    for series in (1, 3, 4):  # west q, north h, south z
        for k in (1, 2):
            d = np.zeros_like(bc0)
            d[:, series] = W[:, k]
            eps = 1e-6

            def J(e):
                A.set_bc_seq(bc0 + e * d)
                return A.cost()

            fd = (J(eps) - J(-eps)) / (2 * eps)
            adj = float(np.sum(G["bc"] * d))
            assert abs(fd - adj) <= 2e-5 * abs(adj) + 1e-10, (series, k, fd, adj)
    A.set_bc_seq(bc0)
    s.close()


def test_gpu_gradient_stride_invariant_and_deterministic(sw, tmp_path):
    g, case, s = _setup(sw, str(tmp_path))
    A, _, _ = _assim(sw, s, case)
    G1 = A.gradient()
    G2 = A.gradient()
    B, _, _ = _assim(sw, s, case, stride=5)
    G3 = B.gradient()
    for key in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(G1[key], G2[key])
        np.testing.assert_array_equal(G1[key], G3[key])
    assert G1["J"] == G2["J"] == G3["J"]
    s.close()


def test_gpu_tiles_and_window_gradient_vs_fd(sw, tmp_path):
    """Tiles restart from their frozen checkpoints and the adjoint is reset at
    each tile top: the gradient is that of the tiled cost, which the FD of the
    tiled cost checks."""
    g, case, s = _setup(sw, str(tmp_path))
    A, _, _ = _assim(sw, s, case, stride=8, n_tiles=3, t_window=4.0)
    phi0 = case.a["phi"].copy()
    G = A.gradient()
    gr = G["phi"]
    idx = np.argsort(-np.abs(np.where(phi0 > 0.05, gr, 0.0)))[:5]
    _check(_fd(lambda x: A.cost(phi=x), phi0, idx), gr[idx], np.max(np.abs(gr)))
    # the tiled gradient is not the untiled one
    B, _, _ = _assim(sw, s, case, stride=8)
    assert not np.array_equal(B.gradient()["phi"], gr)
    s.close()


def test_gpu_frame_cost_and_gradient(sw, tmp_path):
    """Full-frame observations (h, qx, qy per cell): the cost at a single
    observation at t_max equals the misfit of the final state, and its
    gradient passes the FD check."""
    from sw2d_amd import assim
    g, case, s = _setup(sw, str(tmp_path))
    nc = case.n_cell
    rng = np.random.default_rng(3)
    fh = case.a["h0"] + 0.05 * rng.random(nc)
    fqx = 0.01 * rng.standard_normal(nc)
    fqy = 0.01 * rng.standard_normal(nc)
    wq = np.full(nc, 0.5)
    A = assim.Assim(s, [20.0], frames=(fh, fqx, fqy), frame_w=(np.ones(nc), wq))
    A.record()
    J = A.cost()
    h, qx, qy = A.final_state()
    Jref = np.sum((h - fh) ** 2) + np.sum(wq * ((qx - fqx) ** 2 + (qy - fqy) ** 2))
    assert abs(J - Jref) <= 1e-12 * Jref, (J, Jref)
    phi0 = case.a["phi"].copy()
    G = A.gradient()
    gr = G["phi"]
    idx = np.argsort(-np.abs(np.where(phi0 > 0.05, gr, 0.0)))[:5]
    _check(_fd(lambda x: A.cost(phi=x), phi0, idx), gr[idx], np.max(np.abs(gr)))
    s.close()


def test_gpu_twin_experiment_zero_misfit_at_truth(sw, tmp_path):
    """Gauge series sampled by the dry run, used as observations: the cost and
    its gradient vanish at the controls that produced them, and not elsewhere."""
    from sw2d_amd import assim
    g, case, s = _setup(sw, str(tmp_path))
    cells, _ = _gauges(case)
    A = assim.Assim(s, OBS_T, gauge_cells=cells, gauge_h=np.zeros((len(OBS_T), len(cells))))
    A.record()
    series = A.gauge_series()
    assert series.shape == (len(OBS_T), len(cells)) and np.all(series > 0)
    B = assim.Assim(s, OBS_T, gauge_cells=cells, gauge_h=series)
    G = B.gradient()
    assert G["J"] == 0.0
    assert not np.any(G["phi"]) and not np.any(G["K1"]) and not np.any(G["bc"])
    assert B.cost(phi=case.a["phi"] * 0.99) > 0.0
    s.close()


@pytest.mark.parametrize("name", [n for n in e_golden_names() if "muscl" not in n])
def test_gpu_gradient_matches_oracle_reverse_sweep(sw, oracle, tmp_path, name):
    """Every control gradient against the oracle's hand-written reverse sweep
    on the same recorded schedule (checked itself against central FD on the
    CPU): max |gpu - oracle| / max |oracle| <= 1e-10 per control kind.  The
    two differ only in summation order (fixed-order gather on the GPU,
    edge-order scatter in the oracle) and in 1-ulp exp/cbrt differences."""
    g, case, s = _setup(sw, str(tmp_path), name)
    from sw2d_amd import assim
    obs_t = list(g["obs_t"])
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][::5][:8].astype(np.int64)
    obs = np.array([[h0[c] * (1.0 + 0.1 * (k + 1)) for c in cells] for k in range(len(obs_t))])
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs, normalize=True)
    G = A.gradient()
    E = oracle.OracleE(case)
    rec = E.record(obs_t)
    assert len(rec["dt"]) == G["n_steps"]
    O = E.gradient(rec, gauges=cells, obs=obs, normalize=True, n_obs=len(obs_t))
    assert abs(G["J"] - O["J"]) <= TOL * abs(O["J"])
    for key in ("phi", "K1", "K2", "bc"):
        scale = np.max(np.abs(O[key]))
        if scale == 0.0:
            assert not np.any(G[key])
            continue
        e = np.max(np.abs(G[key] - O[key])) / scale
        assert e <= 1e-10, (key, e)
    s.close()
