"""Runoff (model_runoff) and Horton infiltration: case files and fixtures.

The oracle's runoff and infiltration steps are pinned bit for bit by the
compiled reference's maps on the rain_* golden cases (test_oracle_golden.py
runs them): full_rainfall, infiltration_coefficient and runoff_coefficient,
with model_porosity_rainfall everywhere and water_only, with and without
Horton infiltration, rain and evaporation from two stations whose records
start after t = 0 and end before t_max.  Here: the reader refuses what the
reference refuses, and the fixtures are not vacuous (rain with
evaporation, and infiltration, each change the maps, and wet cells dry out
during the runs).
"""
import os

import numpy as np
import pytest

from conftest import regen_case, time_of


def _case(sw, folder, **kw):
    from sw2d_amd import cases
    args = dict(scheme="godunov", friction="strickler", t_max=20.0, dt_map=5.0,
                rain="runoff_coefficient", rain_where="everywhere", infil=True)
    args.update(kw)
    return cases.mixed_case(folder, **args)


def _rewrite(path, fn):
    with open(path) as f:
        lines = f.read().splitlines()
    with open(path, "w") as f:
        f.write("\n".join(fn(lines)) + "\n")


def _set_first(v):
    def f(lines):
        lines[5] = v
        return lines
    return f


@pytest.mark.parametrize("edit,msg", [
    ("coeff", "coefficient < 0 or > 1"),
    ("horton", "parameter is negative"),
    ("code0", "zero or smaller"),
    ("scs", "industrial use"),
    ("h_prop", "industrial use"),
    ("where", "model_porosity_rainfall"),
    ("times", "non-increasing"),
])
def test_case_reader_rejects_what_the_reference_rejects(sw, tmp_path, edit, msg):
    folder = _case(sw, str(tmp_path))
    inp = os.path.join(folder, "input")
    solv = os.path.join(inp, "input.solv")
    if edit == "coeff":
        _rewrite(os.path.join(inp, "runoff_coefficient_map.txt"), _set_first("1.5"))
    elif edit == "horton":
        _rewrite(os.path.join(inp, "infiltration_horton_parameter_map.txt"), _set_first("-1e-4"))
    elif edit == "code0":
        _rewrite(os.path.join(inp, "precipitation_station_codes_map.txt"), _set_first("0"))
    elif edit == "scs":
        _rewrite(solv, lambda l: [x.replace("runoff_coefficient", "scs") for x in l])
    elif edit == "h_prop":
        _rewrite(solv, lambda l: [x.replace("horton", "h_proportional") for x in l])
    elif edit == "where":
        _rewrite(solv, lambda l: [x.replace("everywhere", "somewhere") for x in l])
    elif edit == "times":
        _rewrite(os.path.join(inp, "precipitation_time_series.txt"),
                 lambda l: l[:5] + [l[6], l[5]] + l[7:])
    with pytest.raises(sw._native.SW2DError, match=msg):
        sw.Case(folder)


def test_case_reader_tables(sw, tmp_path):
    c = sw.Case(_case(sw, str(tmp_path)))
    np.testing.assert_array_equal(c["source_models"], [3, 0, 1, 0])   # runoff_coefficient, everywhere, horton, no diffusion
    assert c["rain_t"].shape == (4,) and c["rain_v"].shape == (8,)
    assert set(np.unique(c["rain_code"])) == {1, 2}
    assert np.all((c["runoff_coeff"] >= 0) & (c["runoff_coeff"] <= 1))
    assert np.all(c["infil_rate"] >= 0)


def _final(oracle, case, t_end):
    o = oracle.OracleSolver(case)
    o.run(t_end=t_end)
    return o.state()


@pytest.mark.parametrize("name", ["rain_mixed_godunov_infil_coeff_everywhere",
                                  "rain_mixed_godunov_runoff_coeff_water_only"])
def test_rain_fixtures_are_not_vacuous(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    t_end = time_of(str(g["files"][-1]))
    H = g["h"][-1]
    scale = np.max(np.abs(H))
    for off in ("rain", "infil"):
        case = sw.Case(str(tmp_path))
        m = case.a["source_models"].copy()
        m[0 if off == "rain" else 2] = 0
        case.a["source_models"] = m
        h, _, _ = _final(oracle, case, t_end)
        assert np.max(np.abs(h - H)) > 1e-4 * scale, off
    # wet cells dry out during the run (evaporation, infiltration and the
    # depth clamps take them to exactly zero)
    o = oracle.OracleSolver(sw.Case(str(tmp_path)))
    prev, dried = o.state()[0], 0
    while o.t < t_end:
        o.step(1)
        h = o.state()[0]
        dried += np.count_nonzero((prev > 0) & (h == 0))
        prev = h
    assert dried > 0
