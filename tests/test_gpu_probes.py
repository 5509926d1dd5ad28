"""The drop-in `sw2dSolver_gpu <folder>` writes output/ProbesTimeSeries.txt
as the reference's sw2dSolver does (header of swSp::init, swSp.cpp:754-783;
one line per storage instant from swSp::storeResultProbes,
swSp.cpp:1611-1639), on the reference's functional case sp_toce (mixed
quads and triangles, blending topography, Strickler friction, head loss, q
inflow with critical depth, fr outlet) over 2 s: compared with the file the
compiled reference wrote (tests/golden/make_sp_toce_case.py)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, sp_toce_folder

pytestmark = pytest.mark.gpu


def test_gpu_cli_probe_series_matches_the_reference_on_sp_toce(sw, tmp_path):
    d = str(tmp_path)
    sp_toce_folder(sw, d, probe_variant=True)
    exe = os.path.join(os.path.dirname(sw.__file__), "lib", "sw2dSolver_gpu")
    print("[gpu] sw2dSolver_gpu sp_toce probes", flush=True)
    r = subprocess.run([exe, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = open(os.path.join(d, "output", "ProbesTimeSeries.txt")).read().splitlines()
    ref = str(np.load(os.path.join(GOLDEN, "sp_toce_input.npz"))["probes_txt"]).splitlines()
    assert len(got) == len(ref) == 29
    assert got[:18] == ref[:18]  # header, column names, initial state: text identical
    # the storage instants are identical; values within 1e-12 of each column's scale
    G = np.array([[float(v) for v in l.split()] for l in got[17:]])
    R = np.array([[float(v) for v in l.split()] for l in ref[17:]])
    np.testing.assert_array_equal(G[:, 0], R[:, 0])
    scale = np.maximum(np.abs(R).max(axis=0), 1e-300)
    err = np.abs(G - R).max(axis=0) / scale
    print(f"[gpu] probe columns max rel {err.max():.2e}; identical lines "
          f"{sum(a == b for a, b in zip(got, ref))}/{len(ref)}", flush=True)
    assert err.max() <= 1e-12
