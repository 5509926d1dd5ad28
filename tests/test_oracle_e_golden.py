"""The differentiable-forward oracle (oracle/sp_oracle_e.c) against the
reference's own _E physics (assim/swSpE_hyp.cpp, swSpE_friction.cpp,
swSpE_core.cpp, swSpE_muscl.cpp) compiled from its sources: the dry-run
schedule (dt and observation trigger of every step) and the dry run's final
state must be bit-identical.  With Godunov the frozen replay ends on the same
state; with MUSCL-EVR it does not: the reference's dry run reconstructs with
the previous step's dt (swSpE_checkpoint.cpp:244-251), its replay with the
current one (:658-659), and the replay has to follow the reference there."""
import numpy as np
import pytest

from conftest import e_golden_names, regen_case


@pytest.mark.parametrize("name", e_golden_names())
def test_oracle_e_matches_reference_e_forward(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    E = oracle.OracleE(sw.Case(str(tmp_path)))
    rec = E.record(list(g["obs_t"]), states=True)
    assert len(rec["dt"]) == len(g["dt"])
    assert np.array_equal(rec["dt"], g["dt"])
    assert np.array_equal(rec["trig"], g["trig"])
    h, qx, qy = rec["states"][-1]
    assert np.array_equal(h, g["h"]) and np.array_equal(qx, g["qx"]) and np.array_equal(qy, g["qy"])
    _, (hr, qxr, qyr) = E.replay(rec)
    same = np.array_equal(hr, h) and np.array_equal(qxr, qx) and np.array_equal(qyr, qy)
    assert same == ("muscl" not in name)


def test_e_fixtures_present():
    assert len(e_golden_names()) >= 5
