"""Build the sp_toce fixture: the reference's own functional case for porosity
assimilation, as data (development container only):

    python tests/golden/make_sp_toce.py

Reads /root/reference/configue_fonctionels/sp_toce with the package's case
tooling (geometry.sw2d, maps, boundary series, input.solv) and stores, in
tests/golden/sp_toce.npz:
  - the case arrays and solver parameters (reference cell / edge order);
  - the 41 observation frames reference/ResultHydro_p*.txt in the
    assimilation window [t_max - assimilation_window_duration, t_max]
    (columns water_depth, unit_discharge_x, unit_discharge_y; the values the
    reference's load_single_observation_file reads, assim/swSpE_ad.cpp:123-154);
  - the assimilation settings of input.solv (window, tiles, cost_normalize,
    uniform h/q weights) and the default checkpoint stride (30);
  - the first row of descent_log_reference_-75.8pct.csv: J at phi = 1 and the
    L2 norm, largest entry and its cell of dJ/dphi, as computed by the
    reference (Enzyme adjoint).
"""
import csv
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
SRC = "/root/reference/configue_fonctionels/sp_toce"
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "tests"))


def main():
    from make_golden import load_pkg
    from conftest import time_of
    sw = load_pkg()
    c = sw.Case(SRC)
    solv = {}
    for line in open(os.path.join(SRC, "input", "input.solv")):
        p = line.split("#")[0].split()
        if len(p) >= 2 and not line.startswith(("=", "#")):
            solv[p[0]] = p[1]
    t_max = float(solv["t_max"])
    t_win = t_max - float(solv["assimilation_window_duration"])
    files = sorted(glob.glob(os.path.join(SRC, "reference", "ResultHydro_p*.txt")),
                   key=lambda f: time_of(os.path.basename(f)))
    obs_t, frames = [], []
    for f in files:
        t = time_of(os.path.basename(f))
        if t < t_win - 1e-6 or t > t_max + 1e-6:
            continue
        a = np.loadtxt(f, skiprows=1)
        obs_t.append(t)
        frames.append(a[:, 3:6].T.copy())  # [3, n_cell]: h, qx, qy
    row = next(csv.DictReader(open(os.path.join(SRC, "descent_log_reference_-75.8pct.csv"))))
    expect = dict(J=float(row["J_data"]), grad_l2=float(row["gradJm_L2"]),
                  grad_max=float(row["gradJphys_max"]), grad_argmax=int(row["gradJphys_argmax"]))
    arrays = {"a_" + k: np.asarray(v) for k, v in c.a.items()}
    params = {"p_" + k: v for k, v in c.params.as_dict().items()}
    np.savez_compressed(
        os.path.join(HERE, "sp_toce.npz"), **arrays, **params,
        n_series=c.n_series, n_times=c.n_times,
        obs_t=np.array(obs_t), frames=np.array(frames),
        t_window=t_win, n_tiles=int(solv["n_assimilation_windows"]), stride=30,
        normalize=solv["cost_normalize"] == "true",
        **{"expect_" + k: v for k, v in expect.items()})
    print("sp_toce:", c.n_cell, "cells,", len(obs_t), "frames", expect)


if __name__ == "__main__":
    main()
