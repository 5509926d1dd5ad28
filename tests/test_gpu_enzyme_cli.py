"""The drop-in `sw2d_enzyme_gpu <folder> [mode] [phi] [alpha]` (the reference's
assim/main_enzyme.cpp:33-269) on case folders, through its own input.solv
assimilation keys, checked against the reference's _E physics (the e_*
fixtures made from assim/swSpE_*.cpp) and against the library it drives:

- forward: the maps it writes every dt_map are the _E states; the last one
  equals the reference's _E final state (1e-12; written at 17 decimals);
- gen-ref-e: reference/ maps every dt_map_ad in [t_state_assimilation,
  t_max] at porosity_ref_map, identical to the library's dry-run states, and
  the twin experiment is consistent: the cost at the true porosity is 0;
- taylor: gradient_porosity.csv is the library's gradient, the Taylor test
  of the largest-gradient cell converges (|r - 1| < 0.05 for some eps);
- minimize (lbfgsb, gd, adam, bfgs): descent/descent_log.csv (21 columns of
  swSpE_minimize.cpp:539-566), one phi_iter map per gradient, the cost
  falls, optimized_porosity.txt and alpha_sweep.csv as the reference writes
  them (swSpE_minimize.cpp:1052-1070).
"""
import csv
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import regen_case, rel_err

pytestmark = pytest.mark.gpu
NAME = "e_mixed_godunov_strickler"


def exe(sw):
    return os.path.join(os.path.dirname(sw.__file__), "lib", "sw2d_enzyme_gpu")


def run(sw, folder, *args, cwd=None, timeout=300):
    print("[gpu] sw2d_enzyme_gpu", os.path.basename(folder), *args, flush=True)
    r = subprocess.run([exe(sw), folder, *args], capture_output=True, text=True, timeout=timeout,
                       cwd=cwd or folder)
    return r


def set_keys(folder, **kv):
    p = os.path.join(folder, "input", "input.solv")
    lines = open(p).read().splitlines()
    keep = [l for l in lines if not (l.split() and l.split()[0] in kv)]
    keep += [f"{k}\t{v}" for k, v in kv.items()]
    open(p, "w").write("\n".join(keep) + "\n")


def write_sp_map(path, phi):
    with open(path, "w") as f:
        f.write("Unif\t0\n==== Default Param\nPhi\n1\n==== Distrib\n")
        f.write("".join(f"{float(v)!r}\n" for v in phi))


@pytest.fixture(scope="module")
def twin(sw, tmp_path_factory):
    """The e_mixed case with a true porosity map and references made by
    gen-ref-e: (folder, case, true phi)."""
    d = str(tmp_path_factory.mktemp("enzyme_cli"))
    regen_case(sw, NAME, d)
    case = sw.Case(d)
    phi_true = case.a["phi"].copy()
    x = case.a["cell_x"]
    phi_true[x < 25.0] *= 0.7
    write_sp_map(os.path.join(d, "input", "phi_true.txt"), phi_true)
    set_keys(d, dt_map=5.0, dt_map_ad=1.0, t_state_assimilation=5.0, cost_normalize="true",
             checkpoint_stride=10, porosity_ref_map="input/phi_true.txt", max_it=4,
             minimize_algo="lbfgsb", lbfgsb_max_linesearch=10)
    r = run(sw, d, "gen-ref-e")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return d, sw.Case(d), phi_true


def test_gpu_cli_forward_matches_reference_e(sw, tmp_path):
    d = str(tmp_path)
    g = regen_case(sw, NAME, d)
    set_keys(d, dt_map=5.0)
    r = run(sw, d, "forward")
    assert r.returncode == 0, r.stderr[-2000:]
    maps = sorted(glob.glob(os.path.join(d, "output", "ResultHydro_p*.txt")))
    assert [os.path.basename(m)[-13:] for m in maps] == [
        "00m05s000.txt", "00m10s000.txt", "00m15s000.txt", "00m20s000.txt"]
    m = sw.read_result_map(maps[-1])
    for k, col in (("h", "water_depth"), ("qx", "unit_discharge_x"), ("qy", "unit_discharge_y")):
        assert rel_err(m[col], g[k]) <= 1e-12, (k, rel_err(m[col], g[k]))


def test_gpu_cli_gen_ref_e_twin_is_consistent(sw, twin):
    from sw2d_amd import assim
    d, case, phi_true = twin
    refs = sorted(glob.glob(os.path.join(d, "reference", "ResultHydro_p*.txt")))
    assert len(refs) == 16  # t = 5, 6, ..., 20
    s = sw.SpSolver(case)
    t = [5.0 + k for k in range(16)]
    A = assim.Assim(s, t, frames=[np.zeros((16, case.n_cell))] * 3, stride=10, max_steps=5000)
    A.keep_obs_states(True)
    A.record(phi=phi_true)
    H, QX, QY, hit = A.obs_states()
    assert hit.all()
    for k, p in enumerate(refs):
        m = sw.read_result_map(p)
        np.testing.assert_array_equal(m["water_depth"], H[k])
        np.testing.assert_array_equal(m["unit_discharge_x"], QX[k])
        np.testing.assert_array_equal(m["unit_discharge_y"], QY[k])
    s.close()
    # the cost of the references at the true porosity is exactly 0
    s = sw.SpSolver(case)
    A = assim.Assim(s, t, frames=(H, QX, QY), normalize=True, stride=10, max_steps=5000)
    G = A.gradient(phi=phi_true)
    assert G["J"] == 0.0 and not np.any(G["phi"])
    s.close()


def test_gpu_cli_taylor_gradient_and_ratio(sw, twin, tmp_path):
    from sw2d_amd import assim
    d, case, phi_true = twin
    r = run(sw, d, "taylor", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-2000:]
    rows = list(csv.DictReader(open(os.path.join(d, "gradient_porosity.csv"))))
    assert len(rows) == case.n_cell and list(rows[0].keys()) == ["cell_id", "phi", "gradient"]
    gcli = np.array([float(x["gradient"]) for x in rows])
    # the library's gradient on the same frames (the CLI prints 6 digits)
    refs = sorted(glob.glob(os.path.join(d, "reference", "ResultHydro_p*.txt")))
    fr = [np.array([sw.read_result_map(p)[c] for p in refs])
          for c in ("water_depth", "unit_discharge_x", "unit_discharge_y")]
    s = sw.SpSolver(case)
    A = assim.Assim(s, [5.0 + k for k in range(16)], frames=fr, normalize=True, stride=10,
                    max_steps=5000)
    G = A.gradient(phi=case.a["phi"])
    s.close()
    assert np.max(np.abs(gcli - G["phi"])) <= 1e-5 * np.max(np.abs(G["phi"]))
    tay = list(csv.reader(open(os.path.join(str(tmp_path), "taylor_test_checkpointed.csv"))))
    assert tay[0] == ["cell_id", "grad_enzyme", "epsilon", "J_traj_pert", "fd_grad", "ratio"]
    top = int(np.argmax(np.abs(G["phi"])))
    ratios = [float(row[5]) for row in tay[1:] if row[0] == str(top)]
    assert len(ratios) == 7 and min(abs(x - 1.0) for x in ratios[1:]) < 0.05, ratios
    glob_rows = [row for row in tay[1:] if row[0] == "global"]
    assert len(glob_rows) == 7


DESCENT_COLS = ("iter,J_total,J_data,ridge,rel_init_pct,rel_prev_pct,gradJm_L2,gradJm_Linf,"
                "gradJm_argmax,gradJphys_L2,gradJphys_max,gradJphys_argmax,dh_peak,phi_min,"
                "phi_max,phi_mean,n_phi_lt_0p5,n_phi_lt_0p05,n_phi_upper_clamp,adam_step,"
                "sse_phi_ref").split(",")


@pytest.mark.parametrize("algo", ["lbfgsb", "gd", "adam", "bfgs"])
def test_gpu_cli_minimize_writes_the_reference_outputs(sw, twin, algo):
    d, case, phi_true = twin
    set_keys(d, minimize_algo=algo, max_it=4, adam_step_size=0.02)
    start = os.path.join(d, "input", "phi_start.txt")
    write_sp_map(start, case.a["phi"])
    sweep = os.path.join(d, "alpha_sweep.csv")
    if os.path.exists(sweep):
        os.remove(sweep)
    r = run(sw, d, "minimize", start, "0.0")
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    log = list(csv.reader(open(os.path.join(d, "descent", "descent_log.csv"))))
    assert log[0] == DESCENT_COLS
    rows = log[1:]
    J = [float(x[1]) for x in rows]
    print(f"[gpu] {algo}: {len(rows)} gradients, J {J[0]:.6e} -> {min(J):.6e}", flush=True)
    assert len(rows) >= 2 and min(J) < J[0]
    assert all(re.fullmatch(r"-?\d\.\d{10}e[+-]\d\d", x[1]) for x in rows)
    assert float(rows[0][20]) > 0.0  # sse to the true map
    assert len(glob.glob(os.path.join(d, "descent", "phi_iter", "phi_iter_*.txt"))) == len(rows)
    phi_opt = np.loadtxt(os.path.join(d, "optimized_porosity.txt"), skiprows=5)
    assert phi_opt.shape == (case.n_cell,) and np.all((phi_opt >= 0.0) & (phi_opt <= 1.0))
    sw_rows = list(csv.reader(open(sweep)))
    assert sw_rows[0] == ["alpha", "J_data", "reg_norm2", "J_total", "n_grad"]
    assert int(sw_rows[1][4]) == len(rows)
