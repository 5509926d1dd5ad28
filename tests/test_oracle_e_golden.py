"""The differentiable-forward oracle (oracle/sp_oracle_e.c) against the
reference's own _E physics (assim/swSpE_hyp.cpp, swSpE_friction.cpp,
swSpE_core.cpp) compiled from its sources: the dry-run schedule (dt and
observation trigger of every step) and the final state must be bit-identical."""
import numpy as np
import pytest

from conftest import e_golden_names, regen_case


@pytest.mark.parametrize("name", e_golden_names())
def test_oracle_e_matches_reference_e_forward(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    E = oracle.OracleE(sw.Case(str(tmp_path)))
    rec = E.record(list(g["obs_t"]))
    assert len(rec["dt"]) == len(g["dt"])
    assert np.array_equal(rec["dt"], g["dt"])
    assert np.array_equal(rec["trig"], g["trig"])
    _, (h, qx, qy) = E.replay(rec)
    assert np.array_equal(h, g["h"]) and np.array_equal(qx, g["qx"]) and np.array_equal(qy, g["qy"])


def test_e_fixtures_present():
    assert len(e_golden_names()) >= 3
