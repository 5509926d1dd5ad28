"""Wind source on the GPU (sw2dg_set_wind, model_wind smithbanke).

The two wind_mixed_* golden cases run through the generic map test of
test_gpu_forward.py (1e-12 against the reference's own maps).  Here:

- a table much longer than the fixtures' (8 stations, 400 records between
  t = 0.5 and 6 s, repeated times) on the reduced scale workload, whose 60
  steps run from before the first record to after the last, several records
  per step: every step agrees with the oracle to 1e-12 (the time to 1e-13),
  so the device's bracket search finds the record the reference's sweep
  stops at;
- turning the source off and on again, with a table of another shape, on a
  live context;
- the 4D-Var layer refuses a context with wind (it has no wind source).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _long_table(case, seed=11, n_st=8, n_rec=400, span=(0.5, 6.0)):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(span[0], span[1], n_rec))
    q = n_rec // 4
    t[q:q + 4] = t[q]                             # repeated instants
    vx = rng.uniform(-35.0, 35.0, (n_st, n_rec))
    vy = rng.uniform(-35.0, 35.0, (n_st, n_rec))
    code = (1 + rng.integers(0, n_st, case.n_cell)).astype(np.int32)
    return code, t, vx, vy


def _with_wind(case, code, t, vx, vy):
    case.a["wind_code"], case.a["wind_t"] = code, t
    case.a["wind_x"], case.a["wind_y"] = vx.reshape(-1), vy.reshape(-1)


def test_gpu_long_wind_table_step_by_step_matches_oracle(sw, oracle, tmp_path):
    folder, _ = sw.scale_case(str(tmp_path), px=60, py=60, L=600.0, t_max=3600.0)
    case = sw.Case(folder)
    table = _long_table(case)
    _with_wind(case, *table)
    print("[gpu] wind long table", case.n_cell, "cells", flush=True)
    s = sw.SpSolver(case)
    o = oracle.OracleSolver(case)
    for k in range(60):
        s.step(1)
        o.step(1)
        assert s.clock()["t"] == pytest.approx(o.t, rel=1e-13, abs=0), k
        h, qx, qy = s.state()
        H, QX, QY = o.state()
        assert rel_err(h, H) <= TOL and rel_err(qx, QX) <= TOL and rel_err(qy, QY) <= TOL, k
    assert o.t > table[1][-1]   # the run went through the whole table and past it
    s.close()


def test_gpu_wind_off_and_on_again(sw, oracle, tmp_path):
    folder, _ = sw.scale_case(str(tmp_path), px=40, py=40, L=400.0, t_max=3600.0)
    case = sw.Case(folder)
    print("[gpu] wind toggle", case.n_cell, "cells", flush=True)
    s = sw.SpSolver(case)
    ref = oracle.OracleSolver(case)          # no wind
    code, t, vx, vy = _long_table(case, seed=5, n_st=3, n_rec=50)
    s.set_wind(code, t, vx, vy)
    s.step(20)
    s.set_wind(np.zeros(0, np.int32), None, None, None)   # off: must match the plain solver
    s.set_state(case["h0"], case["qx0"], case["qy0"], 0.0)
    s.step(20)
    ref.step(20)
    for a, b in zip(s.state(), ref.state()):
        assert rel_err(a, b) <= TOL
    code2, t2, vx2, vy2 = _long_table(case, seed=6, n_st=5, n_rec=70)   # another shape
    s.set_wind(code2, t2, vx2, vy2)
    s.set_state(case["h0"], case["qx0"], case["qy0"], 0.0)
    s.step(20)
    _with_wind(case, code2, t2, vx2, vy2)
    o = oracle.OracleSolver(case)
    o.step(20)
    for a, b in zip(s.state(), o.state()):
        assert rel_err(a, b) <= TOL
    assert rel_err(s.state()[1], ref.state()[1]) > 1e-6   # the wind acted
    s.close()


def test_gpu_assim_refuses_wind(sw, tmp_path):
    from sw2d_amd import assim
    folder, _ = sw.scale_case(str(tmp_path), px=20, py=20, L=200.0, t_max=60.0)
    case = sw.Case(folder)
    _with_wind(case, *_long_table(case, n_st=2, n_rec=10))
    s = sw.SpSolver(case)
    with pytest.raises(sw._native.SW2DError, match="has no wind"):
        assim.Assim(s, np.array([10.0, 20.0]), gauge_cells=np.array([0]), gauge_h=np.zeros((2, 1)))
    s.close()


@pytest.mark.parametrize("name", ["wind_dam_small_godunov", "wind_dam_small_muscl_evr"])
def test_gpu_wind_without_transcendentals_is_bitwise(sw, tmp_path, name):
    """Walls only and no friction: nothing on the path calls exp/log, and
    the wind step adds no transcendental on the device (its trigonometry
    runs on the host), so every map equals the reference's bit for bit."""
    from conftest import regen_case, time_of
    g = regen_case(sw, name, str(tmp_path))
    print("[gpu] wind bitwise", name, flush=True)
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


def test_gpu_cli_writes_reference_maps_with_wind(sw, tmp_path):
    """The drop-in `sw2dSolver_gpu <folder>` reads the wind files itself and
    writes the same ResultHydro maps as the reference's `sw2dSolver`."""
    import glob
    import os
    import subprocess
    from conftest import regen_case
    name = "wind_dam_small_muscl_evr"
    g = regen_case(sw, name, str(tmp_path))
    exe = os.path.join(os.path.dirname(sw.__file__), "lib", "sw2dSolver_gpu")
    print("[gpu] cli", name, flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outs = sorted(glob.glob(os.path.join(str(tmp_path), "output", "ResultHydro_p*.txt")))
    assert [os.path.basename(p) for p in outs] == [str(f) for f in g["files"]]
    for k, p in enumerate(outs):
        m = sw.read_result_map(p)
        np.testing.assert_array_equal(m["water_depth"], g["h"][k])
        np.testing.assert_array_equal(m["unit_discharge_x"], g["qx"][k])
        np.testing.assert_array_equal(m["unit_discharge_y"], g["qy"][k])
