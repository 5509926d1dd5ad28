"""Runoff and Horton infiltration on the GPU (sw2dg_set_runoff,
sw2dg_set_infiltration).

The five rain_* golden cases run through the generic map test of
test_gpu_forward.py (1e-12 against the reference's own maps).  Here:

- the walls-only, friction-free rain fixture (full_rainfall everywhere with
  Horton infiltration) bit-identical at every map, also through the drop-in
  CLI: nothing on its path calls exp/log;
- a long rain table (6 stations, 300 records, repeated times, evaporation)
  with every runoff model and both porosity-rainfall modes on the reduced
  scale workload: 40 steps within 1e-12 of the oracle at every step;
- switching the sources off on a live context gives the plain solver;
- the 4D-Var layer refuses a context with runoff or infiltration;
- bad source tables are refused by the C-ABI.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import regen_case, rel_err, time_of

pytestmark = pytest.mark.gpu
TOL = 1e-12


def test_gpu_rain_without_transcendentals_is_bitwise(sw, tmp_path):
    name = "rain_dam_small_godunov_full_infil"
    g = regen_case(sw, name, str(tmp_path))
    print("[gpu] rain bitwise", name, flush=True)
    s = sw.SpSolver(sw.Case(str(tmp_path)))
    for k, f in enumerate(g["files"]):
        t = time_of(str(f))
        if t > 0:
            s.run(t_stop=t)
        h, qx, qy = s.state()
        np.testing.assert_array_equal(h, g["h"][k], err_msg=f"h at {f}")
        np.testing.assert_array_equal(qx, g["qx"][k], err_msg=f"qx at {f}")
        np.testing.assert_array_equal(qy, g["qy"][k], err_msg=f"qy at {f}")
    s.close()
    exe = os.path.join(os.path.dirname(sw.__file__), "lib", "sw2dSolver_gpu")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outs = sorted(glob.glob(os.path.join(str(tmp_path), "output", "ResultHydro_p*.txt")))
    assert len(outs) == len(g["files"])
    for k, p in enumerate(outs):
        m = sw.read_result_map(p)
        np.testing.assert_array_equal(m["water_depth"], g["h"][k])
        np.testing.assert_array_equal(m["unit_discharge_x"], g["qx"][k])


def _rain_table(case, model, water_only, seed=21, n_st=6, n_rec=300):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.3, 4.0, n_rec))
    q = n_rec // 3
    t[q:q + 3] = t[q]
    rate = rng.uniform(-3e-3, 4e-3, (n_st, n_rec))
    code = (1 + rng.integers(0, n_st, case.n_cell)).astype(np.int32)
    coeff = rng.uniform(0.0, 1.0, case.n_cell)
    case.a["source_models"] = np.array([model, int(water_only), 1], np.int32)
    case.a["rain_code"], case.a["rain_t"], case.a["rain_v"] = code, t, rate.reshape(-1)
    case.a["runoff_coeff"] = coeff
    case.a["infil_rate"] = rng.uniform(0.0, 5e-4, case.n_cell)


@pytest.mark.parametrize("model,water_only", [(1, False), (2, False), (3, False), (1, True),
                                              (3, True)])
def test_gpu_long_rain_table_step_by_step_matches_oracle(sw, oracle, tmp_path, model, water_only):
    folder, _ = sw.scale_case(str(tmp_path), px=50, py=50, L=500.0, t_max=3600.0)
    case = sw.Case(folder)
    _rain_table(case, model, water_only)
    print("[gpu] rain long table", model, water_only, case.n_cell, "cells", flush=True)
    s = sw.SpSolver(case)
    o = oracle.OracleSolver(case)
    for k in range(40):
        s.step(1)
        o.step(1)
        assert s.clock()["t"] == pytest.approx(o.t, rel=1e-13, abs=0), k
        h, qx, qy = s.state()
        H, QX, QY = o.state()
        assert rel_err(h, H) <= TOL and rel_err(qx, QX) <= TOL and rel_err(qy, QY) <= TOL, k
    assert o.t > case.a["rain_t"][-1]
    s.close()


def test_gpu_sources_off_give_the_plain_solver(sw, oracle, tmp_path):
    folder, _ = sw.scale_case(str(tmp_path), px=30, py=30, L=300.0, t_max=3600.0)
    case = sw.Case(folder)
    ref = oracle.OracleSolver(case)
    _rain_table(case, 2, False)
    print("[gpu] rain toggle", case.n_cell, "cells", flush=True)
    s = sw.SpSolver(case)
    s.step(10)
    s.set_runoff(0, True)
    s.set_infiltration(0)
    s.set_state(case["h0"], case["qx0"], case["qy0"], 0.0)
    s.step(20)
    ref.step(20)
    for a, b in zip(s.state(), ref.state()):
        assert rel_err(a, b) <= TOL
    s.close()


def test_gpu_assim_refuses_runoff(sw, tmp_path):
    from sw2d_amd import assim
    folder, _ = sw.scale_case(str(tmp_path), px=20, py=20, L=200.0, t_max=60.0)
    case = sw.Case(folder)
    _rain_table(case, 1, True, n_st=2, n_rec=10)
    s = sw.SpSolver(case)
    with pytest.raises(sw._native.SW2DError, match="runoff"):
        assim.Assim(s, np.array([10.0, 20.0]), gauge_cells=np.array([0]), gauge_h=np.zeros((2, 1)))
    s.close()


def test_gpu_source_tables_are_validated(sw, tmp_path):
    """Station codes outside [1, n_station], decreasing record times,
    coefficients outside [0, 1] and negative infiltration rates are refused
    before anything reaches the device."""
    folder, _ = sw.scale_case(str(tmp_path), px=10, py=10, L=100.0, t_max=60.0)
    case = sw.Case(folder)
    print("[gpu] source validation", case.n_cell, "cells", flush=True)
    s = sw.SpSolver(case)
    nc = case.n_cell
    code = np.ones(nc, np.int32)
    t = np.array([0.0, 10.0])
    v = np.zeros((2, 2))
    E = sw._native.SW2DError
    bad_code = code.copy()
    bad_code[3] = 3
    with pytest.raises(E, match="station code"):
        s.set_runoff(1, True, bad_code, t, v)
    with pytest.raises(E, match="non-decreasing"):
        s.set_runoff(1, True, code, t[::-1].copy(), v)
    with pytest.raises(E, match="outside \\[0, 1\\]"):
        s.set_runoff(2, False, code, t, v, np.full(nc, 1.5))
    with pytest.raises(E, match="negative"):
        s.set_infiltration(1, np.full(nc, -1e-4))
    with pytest.raises(E, match="station code"):
        s.set_wind(bad_code, t, v, v)
    s.set_runoff(1, True, code, t, v)   # a valid table is taken
    s.step(2)
    s.close()
