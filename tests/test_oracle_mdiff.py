"""Momentum diffusion (model_momentum_diffusion): the fixtures are not vacuous.

The oracle's parabolic step is pinned bit for bit by the compiled reference's
maps on the mdiff_* golden cases (test_oracle_golden.py runs them).  Its
coefficient is the reference's fixed 1e-6 (swSp.h:244; the input key is never
read), so the effect is small: here, switching it off moves the final maps
by 5e-8 (depth) to 1e-6 (discharge) relative, four to six orders above the
1e-12 parity tolerance.
"""
import numpy as np
import pytest

from conftest import regen_case, rel_err, time_of


@pytest.mark.parametrize("name", ["mdiff_dam_small_godunov", "mdiff_dam_small_muscl_manning_wind"])
def test_momentum_diffusion_fixture_is_not_vacuous(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    m = case.a["source_models"].copy()
    assert m[3] == 1
    m[3] = 0
    case.a["source_models"] = m
    o = oracle.OracleSolver(case)
    o.run(t_end=time_of(str(g["files"][-1])))
    h, qx, qy = o.state()
    assert rel_err(h, g["h"][-1]) > 1e-8
    assert rel_err(qx, g["qx"][-1]) > 1e-7 and rel_err(qy, g["qy"][-1]) > 1e-7


def test_any_value_but_none_turns_it_on(sw, tmp_path):
    """swSolverParam::setModelMomentumDiffusion accepts any value and
    swSp::execute diffuses for anything but none (swSp.cpp:537)."""
    import os
    from sw2d_amd import cases
    folder = cases.dambreak_case(str(tmp_path), nx=6, ny=4, Lx=12.0, Ly=8.0, x_dam=6.0,
                                 scheme="godunov", friction="none", t_max=1.0, mdiff=True)
    solv = os.path.join(folder, "input", "input.solv")
    with open(solv) as f:
        text = f.read()
    with open(solv, "w") as f:
        f.write(text.replace("velocity_independent", "anything"))
    assert sw.Case(folder)["source_models"][3] == 1
