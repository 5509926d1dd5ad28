"""Wind source (model_wind smithbanke): case files, the oracle, known answers.

The oracle's wind step is pinned bit for bit by the compiled reference's maps
on the wind_mixed_* golden cases (test_oracle_golden.py runs them with every
other fixture).  This file checks what those maps leave open:

- the host reader: station codes and the (norm, angle from North) records
  become Cartesian components that match the closed form -norm * (sin a,
  cos a); the reference's fatal input errors are errors here too;
- the record interpolation (swParent::interpolateTimeSeries) against its
  closed forms: clamps at both ends, exact values at records and midpoints,
  numpy.interp inside, duplicate times;
- the fixtures are not vacuous: without the wind the flow differs.
"""
import os

import numpy as np
import pytest

from conftest import regen_case, time_of


def _wind_case(sw, folder):
    from sw2d_amd import cases
    return cases.mixed_case(folder, scheme="godunov", friction="strickler", t_max=20.0, dt_map=5.0,
                            wind=True)


def _read_records(folder):
    rows = np.loadtxt(os.path.join(folder, "input", "wind_time_series.txt"), skiprows=5, ndmin=2)
    return rows[:, 0], rows[:, 1::2].T, rows[:, 2::2].T  # t, norm[st, r], angle[st, r]


def test_case_reader_gives_cartesian_wind(sw, tmp_path):
    folder = _wind_case(sw, str(tmp_path))
    c = sw.Case(folder)
    t, norm, ang = _read_records(folder)
    nt = len(t)
    np.testing.assert_array_equal(c["wind_t"], t)
    assert c["wind_x"].shape == (2 * nt,) and c["wind_y"].shape == (2 * nt,)
    # angle clockwise from North, the vector points where the wind blows to
    a = np.deg2rad(ang)
    np.testing.assert_allclose(c["wind_x"].reshape(2, nt), -norm * np.sin(a), rtol=0, atol=1e-13)
    np.testing.assert_allclose(c["wind_y"].reshape(2, nt), -norm * np.cos(a), rtol=0, atol=1e-13)
    code = c["wind_code"]
    assert code.dtype == np.int32 and set(np.unique(code)) == {1, 2}
    np.testing.assert_array_equal(code, np.where(c["cell_x"] < 30.0, 1, 2))


def _rewrite(path, fn):
    with open(path) as f:
        lines = f.read().splitlines()
    with open(path, "w") as f:
        f.write("\n".join(fn(lines)) + "\n")


@pytest.mark.parametrize("edit,msg", [
    ("code0", "zero or smaller"),
    ("code_frac", "Non-integer"),
    ("code_big", "exceeds the number of stations"),
    ("times", "non-increasing"),
    ("model", "Wind model unknown"),
    ("short_row", "misformatted"),
])
def test_case_reader_rejects_what_the_reference_rejects(sw, tmp_path, edit, msg):
    folder = _wind_case(sw, str(tmp_path))
    inp = os.path.join(folder, "input")
    codes = os.path.join(inp, "wind_station_codes_map.txt")
    series = os.path.join(inp, "wind_time_series.txt")

    def first_code(v):
        def f(lines):
            lines[5] = v
            return lines
        return f

    if edit == "code0":
        _rewrite(codes, first_code("0"))
    elif edit == "code_frac":
        _rewrite(codes, first_code("1.5"))
    elif edit == "code_big":
        _rewrite(codes, first_code("3"))
    elif edit == "times":
        _rewrite(series, lambda l: l[:5] + [l[6], l[5]] + l[7:])
    elif edit == "model":
        _rewrite(os.path.join(inp, "input.solv"),
                 lambda l: [x.replace("smithbanke", "unknown_wind") for x in l])
    elif edit == "short_row":
        _rewrite(series, lambda l: l[:6] + [" ".join(l[6].split()[:-1])] + l[7:])
    with pytest.raises(sw._native.SW2DError, match=msg):
        sw.Case(folder)


def _interp(oracle, t, v, at):
    import ctypes as C
    t = np.ascontiguousarray(t, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    p = C.POINTER(C.c_double)
    return oracle.lib().orc_interp_series(t.ctypes.data_as(p), v.ctypes.data_as(p), len(t), at)


def test_record_interpolation_closed_forms(oracle):
    t = np.array([2.0, 7.0, 11.0, 15.0])
    v = np.array([1.0, -4.0, 8.0, 0.5])
    assert _interp(oracle, t, v, -3.0) == 1.0           # before the first record
    assert _interp(oracle, t, v, 2.0 - 1e-3) == 1.0
    assert _interp(oracle, t, v, 15.0) == 0.5           # at and after the last
    assert _interp(oracle, t, v, 40.0) == 0.5
    for k in range(len(t)):                             # exact at every record
        assert _interp(oracle, t, v, t[k]) == v[k]
    assert _interp(oracle, t, v, 4.5) == -1.5           # midpoints
    assert _interp(oracle, t, v, 9.0) == 2.0
    assert _interp(oracle, t, v, 13.0) == 4.25
    rng = np.random.default_rng(3)
    tt = np.cumsum(rng.uniform(0.1, 3.0, 50))
    vv = rng.standard_normal(50)
    for at in rng.uniform(tt[0], tt[-1], 200):
        assert _interp(oracle, tt, vv, at) == pytest.approx(np.interp(at, tt, vv), rel=1e-13, abs=1e-14)
    # duplicate times: the value jumps at the repeated instant, left value there
    td = np.array([0.0, 5.0, 5.0, 10.0])
    vd = np.array([0.0, 1.0, 3.0, 3.0])
    assert _interp(oracle, td, vd, 5.0) == 1.0
    assert _interp(oracle, td, vd, 2.5) == 0.5
    assert _interp(oracle, td, vd, 7.5) == 3.0
    # one record: its value everywhere
    assert _interp(oracle, np.array([4.0]), np.array([2.5]), 4.0 - 1e-17) == 2.5


def test_wind_fixture_is_not_vacuous(sw, oracle, tmp_path):
    name = "wind_mixed_godunov_strickler"
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    case.a["wind_code"] = np.zeros(0, np.int32)   # same case, wind off
    o = oracle.OracleSolver(case)
    o.run(t_end=time_of(str(g["files"][-1])))
    h, qx, qy = o.state()
    # about 0.5 % of the largest discharge after 20 s: nine orders above the
    # 1e-12 parity tolerance
    scale = np.max(np.abs(g["qx"][-1]))
    assert np.max(np.abs(qx - g["qx"][-1])) > 1e-3 * scale
    assert np.max(np.abs(qy - g["qy"][-1])) > 1e-3 * np.max(np.abs(g["qy"][-1]))


def test_wind_fixture_records_cover_both_clamps(sw, tmp_path):
    """The golden wind cases start before the first record and end after the
    last, so all three branches of the interpolation run in them."""
    for name in ("wind_mixed_godunov_strickler", "wind_mixed_muscl_evr_manning"):
        folder = str(tmp_path / name)
        g = regen_case(sw, name, folder)
        t, _, _ = _read_records(folder)
        assert t[0] > 0.0 and t[-1] < time_of(str(g["files"][-1]))
