"""The CPU oracle against the compiled reference's own outputs (golden fixtures).

The fixtures hold the maps the reference's `sw2dSolver` wrote (17 decimals,
an exact FP64 round trip) on cases rebuilt here byte-for-byte; the oracle
must reproduce them bit for bit at every storage instant.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden, map_time, regen_case


@pytest.mark.parametrize("name", [n for n in golden_names() if not n.endswith("_error")])
def test_oracle_matches_reference_maps(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    o = oracle.OracleSolver(case)
    for k, f in enumerate(g["files"]):
        t = map_time(g, k)
        if t > 0:
            o.run(t_end=t)
        assert abs(o.t - t) <= 1e-9 * t, (o.t, t)
        h, qx, qy = o.state()
        np.testing.assert_array_equal(h, g["h"][k], err_msg=f"{name} h at {f}")
        np.testing.assert_array_equal(qx, g["qx"][k], err_msg=f"{name} qx at {f}")
        np.testing.assert_array_equal(qy, g["qy"][k], err_msg=f"{name} qy at {f}")


def test_oracle_reproduces_reference_fatal(sw, oracle, tmp_path):
    name = "mixed_muscl_closed_cells_error"
    g = regen_case(sw, name, str(tmp_path))
    assert "hypHllc" in str(g["error"])
    o = oracle.OracleSolver(sw.Case(str(tmp_path)))
    with pytest.raises(RuntimeError, match="oracle error 4"):  # ORC_ERR_FLOW_CONFIG
        o.step(1)
