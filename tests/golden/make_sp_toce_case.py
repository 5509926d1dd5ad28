"""Pack the reference's functional case sp_toce as data and record the
reference's probe output on it (development container only):

    python tests/golden/make_sp_toce_case.py

Writes tests/golden/sp_toce_input.npz:
  - every file of configue_fonctionels/sp_toce/input/ as bytes (geometry.sw2d,
    maps, boundary series, input.solv with its assimilation keys, probes),
    so a test can rebuild the case folder where /root/reference is absent;
  - the descent log descent_log_reference_-75.8pct.csv as text;
  - probe golden: the compiled reference sw2dSolver (oracle/_ref/sw2dSolver)
    run on a native variant of the case -- input.solv without the
    assimilation keys (the native parser stops on them), t_max 2 s,
    dt_time_series 0.2 s, and the case's probes_location.txt under the name
    the reference reads (probe_locations.txt, swFile.h:45) without its first
    probe, which lies outside the mesh (the reference then refuses the case,
    swSp.cpp:766-770) -- its output/ProbesTimeSeries.txt (text) and its last
    state map (h, qx, qy at 17 decimals).
"""
import glob
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
SRC = "/root/reference/configue_fonctionels/sp_toce"
REF = os.path.join(REPO, "oracle", "_ref", "sw2dSolver")
sys.path.insert(0, HERE)

# assimilation keys of the reference's AD layer (assim/assimConfig.cpp:27-39)
AD_KEYS = {
    "adam_safe", "adam_step_size", "alpha_l1", "alpha_phi_ref", "alpha_reg", "alpha_spatial",
    "assimilation_window_duration", "assimilation_window_steps", "checkpoint_stride",
    "cost_normalize", "dt_map_ad", "h_weight", "lambda_tanh", "lbfgsb_epsilon", "lbfgsb_m",
    "lbfgsb_max_linesearch", "mask_dry_cells", "mask_wall_gradient", "max_it", "minimize_algo",
    "n_assimilation_windows", "obs_type", "phi_lower_bound", "phi_transform", "phi_upper_bound",
    "porosity_ebauche_map", "porosity_free_zone_map", "porosity_ref_map", "q_weight",
    "t_state_assimilation", "windowed_assimilation", "zonation"}


def native_solv(text, t_max, dt_ts, n_decimal=17):
    out = []
    for line in text.splitlines():
        s = line.strip()
        if not s or s[0] in "#=":
            out.append(line)
            continue
        p = s.split("#")[0].split()
        if not p or p[0].lower() in AD_KEYS or p[0] in ("t_max", "dt_time_series", "n_decimal"):
            continue
        out.append(f"{p[0]} {p[1]}")
    out += [f"t_max {t_max!r}", f"dt_time_series {dt_ts!r}", f"n_decimal {n_decimal}"]
    return "\n".join(out) + "\n"


def probe_variant(folder, t_max=2.0, dt_ts=0.2):
    """The native probe run's case folder (inputs only)."""
    inp = os.path.join(folder, "input")
    with open(os.path.join(inp, "input.solv")) as f:
        txt = f.read()
    with open(os.path.join(inp, "input.solv"), "w") as f:
        f.write(native_solv(txt, t_max, dt_ts))
    with open(os.path.join(inp, "probes_location.txt")) as f:
        lines = f.read().splitlines()
    with open(os.path.join(inp, "probe_locations.txt"), "w") as f:
        f.write("\n".join([lines[0]] + lines[2:]) + "\n")


def main():
    from make_golden import load_pkg
    sw = load_pkg()
    files = {}
    for p in sorted(glob.glob(os.path.join(SRC, "input", "*"))):
        files["in_" + os.path.basename(p)] = np.frombuffer(open(p, "rb").read(), dtype=np.uint8)
    log = open(os.path.join(SRC, "descent_log_reference_-75.8pct.csv")).read()
    tmp = tempfile.mkdtemp(prefix="sw2d_toce_")
    try:
        shutil.copytree(os.path.join(SRC, "input"), os.path.join(tmp, "input"))
        os.chmod(os.path.join(tmp, "input"), 0o755)
        for p in glob.glob(os.path.join(tmp, "input", "*")):
            os.chmod(p, 0o644)
        probe_variant(tmp)
        r = subprocess.run([REF, tmp], capture_output=True, text=True, timeout=1200)
        if r.returncode != 0:
            sys.exit(f"reference failed: {r.stdout[-800:]} {r.stderr[-400:]}")
        probes = open(os.path.join(tmp, "output", "ProbesTimeSeries.txt")).read()
        maps = sorted(glob.glob(os.path.join(tmp, "output", "ResultHydro_p*.txt")))
        last = sw.read_result_map(maps[-1])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(os.path.join(HERE, "sp_toce_input.npz"), descent_log=log, probes_txt=probes,
                        probes_map_name=os.path.basename(maps[-1]), probes_h=last["water_depth"],
                        probes_qx=last["unit_discharge_x"], probes_qy=last["unit_discharge_y"],
                        **files)
    print("sp_toce_input:", len(files), "files;", probes.count("\n"), "probe lines; last map",
          os.path.basename(maps[-1]))


if __name__ == "__main__":
    main()
