"""The reference's functional assimilation case sp_toce on the GPU (4800
cells, quads and triangles, head loss, Strickler friction, q inflow with
critical depth, fr outlet), through the C-ABI.

What can be matched, and how tightly:
- the _E forward is bit-identical to the oracle (itself bit-identical to the
  reference) for the first 100 steps;
- on that horizon the gradient of a frame misfit matches the oracle's reverse
  sweep to 1e-10 for every control kind;
- over the full case (3157 steps) the GPU and glibc exp/log round differently
  in rare cases (the inflow's critical depth exp(log(q/phi/g)/3) first), and
  the wet/dry fronts amplify such 1-ulp differences: changing ONE input of the
  oracle by 1 ulp (a boundary value at one step, or phi of one cell) moves J
  by 0.7e-6 .. 4.2e-6 relative.  The cost at phi = 1 must therefore match the
  reference's descent log (J = 3.1042911536) to 1e-5 relative, not to its 11
  printed digits.  The porosity gradient of this case is ill-conditioned
  (1-ulp changes of one porosity move its L2 norm between 0.09 and 3), so the
  log's gradient figures are not a parity target.  Neither is a finite
  difference over the full window: the frozen, tiled cost is only piecewise
  smooth there (cells cross the wet/dry threshold), and its central
  difference along a random direction does not converge as eps shrinks
  (measured for eps = 1e-3 .. 1e-9).  The adjoint's exactness is shown where
  it is well posed: against the oracle on the bit-identical horizon (and on
  the e_* cases), and against finite differences on the CPU.
"""
import sys
import os

import numpy as np
import pytest

from conftest import load_sp_toce

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.dirname(__file__))


def _assim(sw, s, g, **kw):
    from sw2d_amd import assim
    fr = g["frames"]
    args = dict(frames=(fr[:, 0], fr[:, 1], fr[:, 2]), normalize=bool(g["normalize"]),
                stride=int(g["stride"]), n_tiles=int(g["n_tiles"]), t_window=float(g["t_window"]),
                max_steps=20000)
    args.update(kw)
    return assim.Assim(s, g["obs_t"], **args)


def test_gpu_sp_toce_first_100_steps_bit_identical(sw, oracle):
    from sw2d_amd import assim
    case, g = load_sp_toce(sw)
    print("[gpu] sp_toce first 100 steps", flush=True)
    ones = np.ones(case.n_cell)
    s = sw.SpSolver(case)
    A = assim.Assim(s, g["obs_t"], max_steps=100)
    assert A.record(phi=ones) == 100
    h, qx, qy = A.final_state()
    E = oracle.OracleE(case)
    rec = E.record(g["obs_t"], phi=ones, max_steps=100)
    np.testing.assert_array_equal(A.get_record()["dt"], rec["dt"])
    _, (ho, qxo, qyo) = E.replay(rec, phi=ones)
    np.testing.assert_array_equal(h, ho)
    np.testing.assert_array_equal(qx, qxo)
    np.testing.assert_array_equal(qy, qyo)
    s.close()


def test_gpu_sp_toce_short_horizon_gradient_matches_oracle(sw, oracle):
    from sw2d_amd import assim
    from test_oracle_sp_toce import short_horizon
    case, g = load_sp_toce(sw)
    print("[gpu] sp_toce short-horizon gradient", flush=True)
    ones = np.ones(case.n_cell)
    E = oracle.OracleE(case)
    rec, t_obs, fr = short_horizon(E, g)
    O = E.frame_gradient(rec, fr, phi=ones)
    s = sw.SpSolver(case)
    A = assim.Assim(s, [t_obs], frames=(fr[:, 0], fr[:, 1], fr[:, 2]), max_steps=len(rec["dt"]) + 5)
    G = A.gradient(phi=ones)
    assert G["n_steps"] >= len(rec["dt"])
    assert abs(G["J"] - O["J"]) <= 1e-12 * O["J"], (G["J"], O["J"])
    for key in ("phi", "K1", "K2", "bc"):
        scale = np.max(np.abs(O[key]))
        n = O[key].shape[0]
        e = np.max(np.abs(G[key][:n] - O[key])) / scale
        assert e <= 1e-10, (key, e)
    s.close()


def test_gpu_sp_toce_cost_and_tiled_gradient(sw):
    case, g = load_sp_toce(sw)
    print("[gpu] sp_toce full window", flush=True)
    s = sw.SpSolver(case)
    A = _assim(sw, s, g)
    ones = np.ones(case.n_cell)
    G = A.gradient(phi=ones)
    Jref = float(g["expect_J"])
    print(f"[gpu] J {G['J']:.10e} (log {Jref:.10e})", flush=True)
    assert abs(G["J"] - Jref) <= 1e-5 * Jref, (G["J"], Jref)
    assert G["n_steps"] == 3157
    # the tiled cost on the frozen schedule is the cost the gradient call reports
    assert A.cost(phi=ones) == G["J"]
    gp = G["phi"]
    assert np.all(np.isfinite(gp)) and np.count_nonzero(gp) > case.n_cell // 2
    s.close()


def test_gpu_sp_toce_lbfgsb_descent(sw):
    """The reference's functional case end to end: from phi = 1, data-only
    L-BFGS-B on [0, 1] (memory 6, <= 30 line-search evaluations, 48 tiles,
    window from t = 10 s, truncation window frozen at the start) over the
    GPU adjoint.  The reference's descent log goes 3.104 -> 0.751 (-75.8 %)
    in 9 evaluations; the path is not reproducible digit for digit (the
    case's chaotic sensitivity, a different L-BFGS-B code), the descent is:
    below 0.9 (-71 %) within 16 iterations."""
    from sw2d_amd.porosity import PorosityProblem, run_minimization
    case, g = load_sp_toce(sw)
    print("[gpu] sp_toce L-BFGS-B descent", flush=True)
    s = sw.SpSolver(case)
    A = _assim(sw, s, g)
    P = PorosityProblem(A, np.ones(case.n_cell))
    th, J, tr = run_minimization(P, "lbfgsb", max_iter=16)
    grads = [v for (_, v, k) in tr if k == "grad"]
    print(f"[gpu] J {grads[0]:.6f} -> {J:.6f} in {len(grads)} evaluations", flush=True)
    assert abs(grads[0] - float(g["expect_J"])) <= 1e-5 * float(g["expect_J"])
    assert J < 0.9
    assert np.all((th >= 0.0) & (th <= 1.0))
    s.close()
