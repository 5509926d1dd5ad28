"""MUSCL-EVR _E on the GPU (k_e_slopes + k_e_edges<true>): the dry run
against the reference's own _E physics is in test_gpu_adjoint.py (the
e_*_muscl_* fixtures); here the frozen replay -- which reconstructs with the
step's own dt, as the reference's taped segment does
(swSpE_checkpoint.cpp:658-659), while the dry run used the previous step's --
against the CPU oracle's replay on the recorded schedule: cost within 1e-12
where no transcendental function is on the path (closed boxes without
friction), within 1e-4 where the inflow's exp/log and the friction's cbrt
differ from glibc's by an ulp and the second-order scheme amplifies it
(CHAOTIC in test_gpu_adjoint.py)."""
import numpy as np
import pytest

from conftest import e_golden_names, regen_case, rel_err
from test_gpu_adjoint import CHAOTIC

pytestmark = pytest.mark.gpu
MUSCL = [n for n in e_golden_names() if "muscl" in n]


@pytest.mark.parametrize("name", MUSCL)
def test_gpu_muscl_replay_cost_matches_oracle(sw, oracle, tmp_path, name):
    from sw2d_amd import assim
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    s = sw.SpSolver(case)
    print(f"[gpu] muscl _E replay {name}", flush=True)
    obs_t = list(g["obs_t"])
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][::5][:8].astype(np.int64)
    obs = np.array([[h0[c] * (1.0 + 0.1 * (k + 1)) for c in cells] for k in range(len(obs_t))])
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs)
    A.record()
    rec = A.get_record()
    J = A.cost()
    E = oracle.OracleE(case)
    Jo, (h, qx, qy) = E.replay({"dt": rec["dt"], "bc": rec["bc"], "trig": rec["trig"]}, gauges=cells,
                               obs=obs, n_obs=len(obs_t))
    tol = 1e-4 if name in CHAOTIC else 1e-12
    assert abs(J - Jo) <= tol * abs(Jo), (J, Jo)
    # the gradient refuses MUSCL-EVR until its adjoint is built
    with pytest.raises(Exception):
        A.gradient()
    s.close()
