"""The resident trajectory of the Godunov gradient (sw2dg_assim_set_resident):
the dry run keeps every segment's states, balanced states and boundary
fluxes and the reverse sweep reads them in place of replaying each segment
from its checkpoint.  The gradient must be bit for bit the replaying one:
cost, every control kind, the boundary-value adjoints, the explosion-guard
truncation and the tiled (reset at tile tops) variant; and MUSCL-EVR, whose
replay reconstructs with another dt than its dry run, always replays."""
import numpy as np
import pytest

from conftest import regen_case

pytestmark = pytest.mark.gpu


def _gauges(case, n_obs):
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][::5][:8].astype(np.int64)
    obs = np.array([[h0[c] * (1.0 + 0.1 * (k + 1)) for c in cells] for k in range(n_obs)])
    return cells, obs


def _gradient(sw, case, g, resident, **kw):
    from sw2d_amd import assim
    s = sw.SpSolver(case)
    obs_t = list(g["obs_t"])
    cells, obs = _gauges(case, len(obs_t))
    A = assim.Assim(s, obs_t, gauge_cells=cells, gauge_h=obs, normalize=True, resident=resident, **kw)
    G = A.gradient()
    out = dict(G, resident=A.resident_segments(), trunc=A.truncation(), peak=A.adjoint_peak())
    s.close()
    return out


@pytest.mark.parametrize("name,kw", [
    ("e_mixed_godunov_strickler", dict(stride=7)),
    ("e_scale_godunov_strickler", dict(stride=11, n_tiles=3)),
    ("e_mixed_godunov_manning_nohl", dict(stride=5, explosion_ratio=10.0)),
])
def test_gpu_resident_gradient_is_the_replaying_one(sw, tmp_path, name, kw):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    print(f"[gpu] resident vs replay {name} {kw}", flush=True)
    on = _gradient(sw, case, g, True, **kw)
    off = _gradient(sw, case, g, False, **kw)
    print(f"[gpu]   {on['resident']} resident segments, truncation {on['trunc']}", flush=True)
    assert on["resident"] > 0 and off["resident"] == 0
    assert on["J"] == off["J"] and on["n_steps"] == off["n_steps"]
    assert on["trunc"] == off["trunc"] and on["peak"] == off["peak"]
    for key in ("phi", "K1", "K2", "bc"):
        np.testing.assert_array_equal(on[key], off[key])


def test_gpu_resident_is_off_for_muscl(sw, tmp_path):
    g = regen_case(sw, "e_mixed_muscl_strickler", str(tmp_path))
    case = sw.Case(str(tmp_path))
    print("[gpu] resident with MUSCL-EVR", flush=True)
    assert _gradient(sw, case, g, True)["resident"] == 0
