"""MUSCL-EVR _E on the GPU: forward (k_e_slopes, k_e_edges_muscl, k_e_update) and its
adjoint (k_r_edge<true> + k_r_slopes + k_r_bnd<true>).

The dry run against the reference's own _E physics is in test_gpu_adjoint.py
(the e_*_muscl_* fixtures).  Here:
- the frozen replay -- which reconstructs with the step's own dt, as the
  reference's taped segment does (swSpE_checkpoint.cpp:658-659), while the
  dry run used the previous step's -- against the CPU oracle's replay on the
  recorded schedule, the oracle built with the kernels' exp / log / cbrt:
  cost and final state bit-identical over the full run, and within 1e-12 of
  the glibc oracle (the reference's arithmetic);
- the gradient against the oracle's reverse sweep (itself checked against
  central differences on the CPU, slope limiter included) on the same
  recorded schedule over the full run, every control kind;
- the gradient against central differences of the GPU's own replay cost
  (one segment covers the window: the gradient is then that of the
  continuous replay the cost runs);
- over several segments, each replayed from its dry-run checkpoint, against
  the oracle with the same restarts.
"""
import numpy as np
import pytest

from conftest import e_golden_names, regen_case

pytestmark = pytest.mark.gpu
MUSCL = [n for n in e_golden_names() if "muscl" in n]


def _gauges(case, n_obs):
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][::5][:8].astype(np.int64)
    obs = np.array([[h0[c] * (1.0 + 0.1 * (k + 1)) for c in cells] for k in range(n_obs)])
    return cells, obs


def _open(sw, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    return g, case, sw.SpSolver(case)


@pytest.mark.parametrize("name", MUSCL)
def test_gpu_muscl_replay_cost_matches_oracle(sw, oracle, tmp_path, name):
    from sw2d_amd import assim
    g, case, s = _open(sw, tmp_path, name)
    print(f"[gpu] muscl _E replay {name}", flush=True)
    obs_t = list(g["obs_t"])
    cells, obs = _gauges(case, len(obs_t))
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs)
    A.record()
    rec = A.get_record()
    J = A.cost()
    rr = {"dt": rec["dt"], "bc": rec["bc"], "trig": rec["trig"]}
    Jd, _ = oracle.OracleE(case, math="device").replay(rr, gauges=cells, obs=obs, n_obs=len(obs_t))
    assert J == Jd, (J, Jd)
    Jo, _ = oracle.OracleE(case).replay(rr, gauges=cells, obs=obs, n_obs=len(obs_t))
    assert abs(J - Jo) <= 1e-12 * abs(Jo), (J, Jo)
    s.close()


@pytest.mark.parametrize("name", MUSCL)
def test_gpu_muscl_gradient_matches_oracle_reverse_sweep(sw, oracle, tmp_path, name):
    """Over the whole run: max |gpu - oracle| / max |oracle| <= 1e-10 per
    control kind (the two reverse sweeps sum in different orders), and the
    same non-finite entries where the adjoint overflows.  The
    oracle's sweep has no explosion guard, so the guard is off here (on
    e_scale_muscl_nofriction the adjoint of h grows past the default 1e8
    ratio and the guarded sweep stops early; the guard has its own tests).
    One segment spans the run, so that the forward recompute is the oracle's
    continuous replay: restarts from dry-run checkpoints, which
    reconstruct with the previous step's dt, differ from it by ~1e-3 of J on
    that case, and are checked against the oracle with the same restarts
    below."""
    from sw2d_amd import assim
    g, case, s = _open(sw, tmp_path, name)
    print(f"[gpu] muscl gradient vs oracle {name}", flush=True)
    obs_t = list(g["obs_t"])
    cells, obs = _gauges(case, len(obs_t))
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs)
    n_steps = A.record()
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs, normalize=True, explosion_ratio=0.0,
                    stride=n_steps)
    G = A.gradient()
    assert A.truncation() == (0, 0)  # the whole run was swept
    rec = A.get_record()
    E = oracle.OracleE(case, math="device")
    O = E.gradient({"dt": rec["dt"], "bc": rec["bc"], "trig": rec["trig"]}, gauges=cells, obs=obs,
                   normalize=True, n_obs=len(obs_t))
    print(f"[gpu]   J {G['J']!r} oracle {O['J']!r}", flush=True)
    assert abs(G["J"] - O["J"]) <= 1e-12 * abs(O["J"])
    for key in ("phi", "K1", "K2", "bc"):
        # where the adjoint overflows (e_scale_muscl_nofriction: 0/0 and inf
        # reach most cells), the cell controls are returned as 0, the
        # reference's rule (swSpE_checkpoint.cpp:1197-1220); the boundary
        # gradient is returned as computed
        ok = np.isfinite(O[key])
        if not ok.all():
            print(f"[gpu]   {key}: {np.count_nonzero(~ok)} non-finite oracle entries", flush=True)
            if key == "bc":
                assert not np.isfinite(G[key][~ok]).any()
            else:
                assert not np.any(G[key][~ok])
        o, gk = O[key][ok], G[key][ok]
        scale = np.max(np.abs(o)) if o.size else 0.0
        if scale == 0.0:
            assert not np.any(gk)
            continue
        e = np.max(np.abs(gk - o)) / scale
        print(f"[gpu]   {key} rel {e:.3e}", flush=True)
        assert e <= 1e-10, (key, e)
    s.close()


def test_gpu_muscl_gradient_vs_central_fd(sw, tmp_path):
    """<dJ/dx, v> against central differences of the GPU replay cost on the
    recorded schedule: porosity, both Strickler coefficients and each
    boundary series (q inflow, fr outlet, h and z levels)."""
    from sw2d_amd import assim
    g, case, s = _open(sw, tmp_path, "e_mixed_muscl_strickler")
    print("[gpu] muscl gradient vs fd", flush=True)
    obs_t = list(g["obs_t"])
    cells, obs = _gauges(case, len(obs_t))
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs, normalize=True)
    G = A.gradient()
    # central differences carry ~ 1e-16 |J| / e of round-off
    floor = lambda e: 4e-16 * abs(G["J"]) / e
    rng = np.random.default_rng(3)
    base = {"phi": case.a["phi"], "K1": case.a["fc1"], "K2": case.a["fc2"]}
    open_ = case.a["phi"] > 0.05
    for key, x0 in base.items():
        v = np.where(open_, rng.standard_normal(case.n_cell), 0.0) * np.maximum(np.abs(x0), 1e-3)
        e = 1e-7
        fd = (A.cost(**{key: x0 + e * v}) - A.cost(**{key: x0 - e * v})) / (2 * e)
        ad = float(np.dot(G[key], v))
        print(f"[gpu]   {key} fd {fd!r} adj {ad!r}", flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad) + floor(e), (key, fd, ad)
    bc0 = A.get_record()["bc"].copy()
    for series in (1, 2, 3, 4):
        v = np.zeros_like(bc0)
        v[:, series] = np.linspace(0.0, 1.0, len(bc0))
        e = 1e-6
        A.set_bc_seq(bc0 + e * v)
        Jp = A.cost()
        A.set_bc_seq(bc0 - e * v)
        Jm = A.cost()
        A.set_bc_seq(bc0)
        fd = (Jp - Jm) / (2 * e)
        ad = float(np.sum(G["bc"] * v))
        print(f"[gpu]   bc{series} fd {fd!r} adj {ad!r}", flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad) + floor(e), (series, fd, ad)
    s.close()


@pytest.mark.parametrize("name", ["e_dambreak_muscl", "e_mixed_muscl_strickler"])
def test_gpu_muscl_gradient_segments_restart_from_dry_run_states(sw, oracle, tmp_path, name):
    """With several segments each one is replayed from its dry-run checkpoint
    (reconstructed with the previous step's dt), so the MUSCL-EVR gradient
    depends on the stride, as the reference's does: stride 5 against the
    oracle's reverse sweep with the same restarts (orc_e_set_checkpoints, the
    dry run's states on the GPU's schedule); and deterministic."""
    from sw2d_amd import assim
    g, case, s = _open(sw, tmp_path, name)
    print(f"[gpu] muscl gradient over segments {name}", flush=True)
    obs_t = list(g["obs_t"])
    cells, obs = _gauges(case, len(obs_t))
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs, normalize=True, stride=5)
    G = A.gradient()
    G2 = A.gradient()
    for key in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(G[key], G2[key])
    rec = A.get_record()
    rec = {"dt": rec["dt"], "bc": rec["bc"], "trig": rec["trig"]}
    E = oracle.OracleE(case, math="device")
    S = E.dry_states(rec)
    O = E.gradient(rec, gauges=cells, obs=obs, normalize=True, n_obs=len(obs_t), checkpoints=(5, S))
    O1 = E.gradient(rec, gauges=cells, obs=obs, normalize=True, n_obs=len(obs_t))
    print(f"[gpu]   J {G['J']!r} oracle {O['J']!r} (one segment {O1['J']!r})", flush=True)
    assert abs(G["J"] - O["J"]) <= 1e-12 * abs(O["J"])
    for key in ("phi", "K1", "K2", "bc"):
        scale = np.max(np.abs(O[key])) if O[key].size else 0.0
        if scale == 0.0:
            continue
        e = np.max(np.abs(G[key] - O[key])) / scale
        moved = np.max(np.abs(O1[key] - O[key])) / scale
        print(f"[gpu]   {key} rel {e:.3e} (restarts move it by {moved:.3e})", flush=True)
        assert e <= 1e-10, (key, e)
    s.close()
