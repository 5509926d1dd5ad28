"""Momentum diffusion on the GPU (sw2dg_set_momentum_diffusion).

The mdiff_* golden cases run through the generic map test of
test_gpu_forward.py (1e-12 against the reference's own maps).  Here: the
friction-free fixture is bit-identical at every map (no exp/log on its
path), also through the drop-in CLI; switching it off on a live context
gives the plain solver; the 4D-Var layer refuses it.  On the sub-micron
mdiff_micro_* fixtures steps take up to six explicit substeps (the device
time loop leaves them to the host, which runs them and resumes): bitwise the
reference's final map and the oracle step by step.  A dry cell stops the run
with the reference's NaN-flux error.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import map_time, regen_case, rel_err, time_of

pytestmark = pytest.mark.gpu


def test_gpu_momentum_diffusion_is_bitwise_and_cli(sw, tmp_path):
    name = "mdiff_dam_small_godunov"
    g = regen_case(sw, name, str(tmp_path))
    print("[gpu] mdiff bitwise", name, flush=True)
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
        np.testing.assert_array_equal(sw.read_result_map(p)["unit_discharge_y"], g["qy"][k])


def test_gpu_momentum_diffusion_off_and_assim(sw, oracle, tmp_path):
    from sw2d_amd import assim
    name = "mdiff_dam_small_muscl_manning_wind"
    regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    print("[gpu] mdiff toggle", name, flush=True)
    s = sw.SpSolver(case)
    with pytest.raises(sw._native.SW2DError, match="momentum diffusion|has no"):
        assim.Assim(s, np.array([1.0]), gauge_cells=np.array([0]), gauge_h=np.zeros((1, 1)))
    s.set_momentum_diffusion(0)
    s.set_state(case["h0"], case["qx0"], case["qy0"], 0.0)
    s.step(30)
    m = case.a["source_models"].copy()
    m[3] = 0
    case.a["source_models"] = m
    o = oracle.OracleSolver(case)
    o.step(30)
    for a, b in zip(s.state(), o.state()):
        assert rel_err(a, b) <= 1e-12
    s.close()


@pytest.mark.parametrize("name", ["mdiff_micro_substeps3_godunov", "mdiff_micro_substeps2_godunov"])
def test_gpu_several_substeps_bitwise(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    print("[gpu] mdiff substeps", name, flush=True)
    # one step per call: every pending step ends a launch
    s = sw.SpSolver(case)
    o = oracle.OracleSolver(case)
    for k in range(12):
        s.step(1)
        o.step(1)
        for a, b in zip(s.state(), o.state()):
            np.testing.assert_array_equal(a, b, err_msg=f"step {k + 1}")
    # the whole run (32-step launches) and the reference's final map
    s.run(t_stop=map_time(g, 0))
    h, qx, qy = s.state()
    np.testing.assert_array_equal(h, g["h"][0])
    np.testing.assert_array_equal(qx, g["qx"][0])
    np.testing.assert_array_equal(qy, g["qy"][0])
    s.close()
    exe = os.path.join(os.path.dirname(sw.__file__), "lib", "sw2dSolver_gpu")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outs = sorted(glob.glob(os.path.join(str(tmp_path), "output", "ResultHydro_p*.txt")))
    np.testing.assert_array_equal(sw.read_result_map(outs[-1])["unit_discharge_x"], g["qx"][0])


def test_gpu_dry_cell_stops_like_the_reference(sw, tmp_path):
    g = regen_case(sw, "mdiff_dry_cell_error", str(tmp_path))
    print("[gpu] mdiff dry cell", flush=True)
    s = sw.SpSolver(sw.Case(str(tmp_path)))
    s.step(1)
    assert np.isnan(s.state()[1]).sum() >= 100
    with pytest.raises(sw._native.SW2DError, match="Not a number"):
        s.step(1)
    s.close()
