"""Generate the golden fixtures of the SP forward step from the compiled reference.

Run in the development container (needs oracle/_ref/sw2dSolver, built from
/root/reference by `make -C oracle ref`):

    python tests/golden/make_golden.py [--only NAME ...]

For every case below the script writes the case folder with the package's
case tooling (deterministic: fixed seeds), runs the reference's own
`sw2dSolver <folder>` on it, and stores the state maps it writes
(ResultHydro_p*.txt, 17 decimals = exact FP64 round trip) in
tests/golden/<name>.npz together with SHA-256 digests of the input files, so
a test can regenerate the same inputs and check it did.  Cases where the
reference stops on a fatal error record that error instead.

The e_* fixtures hold the differentiable forward of the 4D-Var layer
(dt / observation-trigger schedule of the dry run and the final state),
produced by oracle/_ref/ref_e_driver from the reference's assim/swSpE_*.cpp.
"""
import glob
import hashlib
import importlib.util
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
PKG = os.path.join(REPO, "inria-internship-project-sw2d-enzyme_amd")
REF = os.path.join(REPO, "oracle", "_ref", "sw2dSolver")
REF_E = os.path.join(REPO, "oracle", "_ref", "ref_e_driver")
OUT = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    if "sw2d_amd" in sys.modules:
        return sys.modules["sw2d_amd"]
    spec = importlib.util.spec_from_file_location("sw2d_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    m = importlib.util.module_from_spec(spec)
    sys.modules["sw2d_amd"] = m
    spec.loader.exec_module(m)
    return m


# name -> (generator, kwargs).  Shared with the tests through golden_cases().
def golden_cases():
    return {
        # polygon cells: interior bricks have 6 internal edges
        "bricks_muscl_evr_strickler": ("mixed_case", dict(px=14, py=12, mesh="bricks",
                                                           scheme="muscl_evr", friction="strickler",
                                                           closed_fraction=0.0, t_max=20.0,
                                                           dt_map=5.0)),
        "bricks_godunov_manning": ("mixed_case", dict(px=14, py=12, mesh="bricks", scheme="godunov",
                                                      friction="manning", t_max=20.0, dt_map=5.0)),
        "dam_small_godunov": ("dambreak_case", dict(nx=20, ny=10, Lx=40.0, Ly=20.0, x_dam=20.0,
                                                     scheme="godunov", friction="none", t_max=3.0,
                                                     dt_map=1.0)),
        "dam_small_muscl_manning": ("dambreak_case", dict(nx=20, ny=10, Lx=40.0, Ly=20.0,
                                                          x_dam=20.0, scheme="muscl_evr",
                                                          friction="manning", t_max=3.0,
                                                          dt_map=1.0)),
        "dambreak_20k_muscl_manning": ("dambreak_case", dict(scheme="muscl_evr",
                                                             friction="manning", t_max=10.0)),
        **{f"mixed_{s}_{t}_{f}": ("mixed_case", dict(scheme=s, topo=t, friction=f, t_max=20.0,
                                                     dt_map=5.0,
                                                     closed_fraction=0.0 if s == "muscl_evr" else 0.04))
           for s in ("godunov", "muscl_evr") for t in ("discontinuous", "blending")
           for f in ("strickler", "manning")},
        # Chezy friction (swSourceFriction.cpp, exp(-g |u| / C^2 / h dt))
        **{f"mixed_{s}_discontinuous_chezy": ("mixed_case", dict(
            scheme=s, friction="chezy", t_max=20.0, dt_map=5.0,
            closed_fraction=0.0 if s == "muscl_evr" else 0.04)) for s in ("godunov", "muscl_evr")},
        # wind source, model_wind smithbanke (swSpSourceWind.cpp); the dam_small
        # ones have walls only and no friction: no exp/log on their path
        "wind_dam_small_godunov": ("dambreak_case", dict(nx=20, ny=10, Lx=40.0, Ly=20.0, x_dam=20.0,
                                                         scheme="godunov", friction="none",
                                                         t_max=3.0, dt_map=1.0, wind=True)),
        "wind_dam_small_muscl_evr": ("dambreak_case", dict(nx=20, ny=10, Lx=40.0, Ly=20.0,
                                                           x_dam=20.0, scheme="muscl_evr",
                                                           friction="none", t_max=3.0, dt_map=1.0,
                                                           wind=True)),
        "wind_mixed_godunov_strickler": ("mixed_case", dict(scheme="godunov", friction="strickler",
                                                            t_max=20.0, dt_map=5.0, wind=True)),
        "wind_mixed_muscl_evr_manning": ("mixed_case", dict(scheme="muscl_evr",
                                                            friction="manning", t_max=20.0,
                                                            dt_map=5.0, closed_fraction=0.0,
                                                            wind=True)),
        # runoff (rain and evaporation) and Horton infiltration
        # (swSpSourceRunoffInfiltration.cpp), both porosity-rainfall modes
        "rain_dam_small_godunov_full_infil": ("dambreak_case", dict(
            nx=20, ny=10, Lx=40.0, Ly=20.0, x_dam=20.0, scheme="godunov", friction="none",
            t_max=3.0, dt_map=1.0, rain="full_rainfall", rain_where="everywhere", infil=True)),
        "rain_mixed_godunov_infil_coeff_everywhere": ("mixed_case", dict(
            scheme="godunov", friction="strickler", t_max=20.0, dt_map=5.0,
            rain="infiltration_coefficient", rain_where="everywhere", infil=True)),
        "rain_mixed_godunov_runoff_coeff_everywhere": ("mixed_case", dict(
            scheme="godunov", friction="manning", t_max=20.0, dt_map=5.0,
            rain="runoff_coefficient", rain_where="everywhere")),
        "rain_mixed_godunov_runoff_coeff_water_only": ("mixed_case", dict(
            scheme="godunov", friction="strickler", t_max=20.0, dt_map=5.0, closed_fraction=0.0,
            rain="runoff_coefficient", rain_where="water_only", infil=True)),
        "rain_mixed_muscl_evr_full_water_only": ("mixed_case", dict(
            scheme="muscl_evr", friction="manning", t_max=20.0, dt_map=5.0, closed_fraction=0.0,
            rain="full_rainfall", rain_where="water_only", infil=True, wind=True)),
        # momentum diffusion (swSpParab.cpp); every cell wet, since the
        # reference divides q by h of both cells of each edge
        "mdiff_dam_small_godunov": ("dambreak_case", dict(
            nx=20, ny=10, Lx=40.0, Ly=20.0, x_dam=20.0, scheme="godunov", friction="none",
            t_max=3.0, dt_map=1.0, mdiff=True)),
        "mdiff_dam_small_muscl_manning_wind": ("dambreak_case", dict(
            nx=20, ny=10, Lx=40.0, Ly=20.0, x_dam=20.0, scheme="muscl_evr", friction="manning",
            t_max=3.0, dt_map=1.0, mdiff=True, wind=True)),
        # momentum diffusion with several explicit substeps per step: the
        # bound area phi / (2 D) falls below dt / 2 only on sub-micron cells
        # (about 3 and 2 substeps at the start of these two); the map names
        # carry milliseconds, so the one map written is the final state
        "mdiff_micro_substeps3_godunov": ("dambreak_case", dict(
            nx=20, ny=10, Lx=1.0e-6, Ly=5.0e-7, x_dam=5.0e-7, scheme="godunov", friction="none",
            t_max=1.0e-7, mdiff=True)),
        "mdiff_micro_substeps2_godunov": ("dambreak_case", dict(
            nx=20, ny=10, Lx=1.3e-6, Ly=6.5e-7, x_dam=6.5e-7, scheme="godunov", friction="none",
            t_max=1.3e-7, mdiff=True)),
        # momentum diffusion divides q by h of both cells of every edge
        # (swSpParab.cpp:74-75): a dry cell gives 0 / 0, and the reference
        # stops at the next flux computation (swSpHyp.cpp:391)
        "mdiff_dry_cell_error": ("dambreak_case", dict(
            nx=20, ny=10, Lx=40.0, Ly=20.0, x_dam=20.0, h_right=0.0, scheme="godunov",
            friction="none", t_max=3.0, dt_map=1.0, mdiff=True)),
        # the reference stops on an impossible HLLC regime (swSpHyp.cpp:729-739)
        "mixed_muscl_closed_cells_error": ("mixed_case", dict(scheme="muscl_evr", friction="strickler",
                                                             t_max=20.0, closed_fraction=0.04)),
    }


# Differentiable (_E) forward: name -> (generator, kwargs, observation times).
# The reference's _E physics (assim/swSpE_*.cpp) runs through
# oracle/ref_e_driver.cpp, which prints dt and the observation trigger of
# every dry-run step and the final state.
def e_golden_cases():
    return {
        "e_scale_godunov_strickler": ("scale_case", dict(px=30, py=30, L=300.0, scheme="godunov",
                                                         t_max=120.0, friction="strickler",
                                                         manning=30.0), [30.0, 60.0, 90.0, 120.0]),
        "e_mixed_godunov_strickler": ("mixed_case", dict(scheme="godunov", friction="strickler",
                                                         t_max=20.0), [5.0, 10.0, 15.0, 20.0]),
        "e_mixed_godunov_manning_nohl": ("mixed_case", dict(scheme="godunov", friction="manning",
                                                            t_max=20.0, head_loss=False),
                                         [5.0, 10.0, 15.0, 20.0]),
        "e_bricks_godunov_strickler": ("mixed_case", dict(px=14, py=12, mesh="bricks",
                                                          scheme="godunov", friction="strickler",
                                                          t_max=20.0), [5.0, 10.0, 15.0, 20.0]),
        "e_bricks_muscl_strickler": ("mixed_case", dict(px=14, py=12, mesh="bricks",
                                                        scheme="muscl_evr", friction="strickler",
                                                        closed_fraction=0.0, t_max=20.0),
                                     [5.0, 10.0, 15.0, 20.0]),
        "e_scale_muscl_strickler": ("scale_case", dict(px=30, py=30, L=300.0, scheme="muscl_evr",
                                                       t_max=120.0, friction="strickler",
                                                       manning=30.0), [30.0, 60.0, 90.0, 120.0]),
        "e_mixed_muscl_strickler": ("mixed_case", dict(scheme="muscl_evr", friction="strickler",
                                                       t_max=20.0), [5.0, 10.0, 15.0, 20.0]),
        "e_scale_muscl_nofriction": ("scale_case", dict(px=30, py=30, L=300.0, scheme="muscl_evr",
                                                        t_max=120.0, friction="none"),
                                     [30.0, 60.0, 90.0, 120.0]),
        # closed boxes, no friction: no transcendental function on the path
        # (walls give no critical-depth inflow), so the GPU can match exactly
        "e_dambreak_muscl": ("dambreak_case", dict(nx=40, ny=20, Lx=80.0, Ly=40.0, x_dam=40.0,
                                                   scheme="muscl_evr", friction="none", t_max=10.0),
                             [2.5, 5.0, 7.5, 10.0]),
        "e_drybreak_muscl": ("dambreak_case", dict(nx=40, ny=20, Lx=80.0, Ly=40.0, x_dam=30.0,
                                                   h_right=0.0, scheme="muscl_evr", friction="none",
                                                   t_max=10.0), [2.5, 5.0, 7.5, 10.0]),
    }


def run_ref_e(folder, obs):
    r = subprocess.run([REF_E, folder, str(len(obs))] + [repr(float(x)) for x in obs],
                       capture_output=True, text=True, timeout=1200)
    if r.returncode != 0:
        raise RuntimeError(f"ref_e_driver failed ({r.returncode}): {r.stderr[-400:]}")
    dt, trig, S = [], [], []
    for line in r.stdout.splitlines():
        p = line.split()
        if p and p[0] == "step":
            dt.append(float(p[2]))
            trig.append(int(p[3]))
        elif p and p[0] == "cell":
            S.append([float(p[2]), float(p[3]), float(p[4])])
    S = np.array(S)
    return np.array(dt), np.array(trig, dtype=np.int32), S[:, 0], S[:, 1], S[:, 2]


def main_e(sw, only=None):
    if not os.path.exists(REF_E):
        sys.exit(f"{REF_E} missing: run `make -C {os.path.join(REPO, 'oracle')} ref` first")
    for name, (gen, kw, obs) in e_golden_cases().items():
        if only and name not in only:
            continue
        tmp = tempfile.mkdtemp(prefix="sw2d_golden_e_")
        try:
            make_case(sw, tmp, gen, kw)
            dig = input_digests(tmp)
            dt, trig, h, qx, qy = run_ref_e(tmp, obs)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), obs_t=np.array(obs), dt=dt,
                                trig=trig, h=h, qx=qx, qy=qy, name=name, generator=gen,
                                kwargs=repr(kw), digest_names=np.array(list(dig.keys())),
                                digest_values=np.array(list(dig.values())))
            print(name, "steps:", len(dt))
        finally:
            shutil.rmtree(tmp, ignore_errors=True)


def make_case(sw, folder, gen, kw):
    from sw2d_amd import cases
    getattr(cases, gen)(folder, **kw)


def input_digests(folder):
    d = {}
    for p in sorted(glob.glob(os.path.join(folder, "input", "*"))):
        with open(p, "rb") as f:
            d[os.path.basename(p)] = hashlib.sha256(f.read()).hexdigest()
    return d


def main(only=None):
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C {os.path.join(REPO, 'oracle')} ref` first")
    sw = load_pkg()
    for name, (gen, kw) in golden_cases().items():
        if only and name not in only:
            continue
        tmp = tempfile.mkdtemp(prefix="sw2d_golden_")
        try:
            make_case(sw, tmp, gen, kw)
            r = subprocess.run([REF, tmp], capture_output=True, text=True, timeout=1200)
            dig = input_digests(tmp)
            meta = dict(name=name, generator=gen, kwargs=repr(kw),
                        digest_names=np.array(list(dig.keys())),
                        digest_values=np.array(list(dig.values())))
            if r.returncode != 0:
                log = open(os.path.join(tmp, "log", "console_Solver.log")).read()
                fatal = [l for l in log.splitlines() if "[Fat.]" in l]
                np.savez_compressed(os.path.join(OUT, name + ".npz"), error=fatal[-1] if fatal else "fatal",
                                    **meta)
                print(name, "reference error:", fatal[-1] if fatal else r.returncode)
                continue
            outs = sorted(glob.glob(os.path.join(tmp, "output", "ResultHydro_p*.txt")))
            times, H, QX, QY = [], [], [], []
            for p in outs:
                m = sw.read_result_map(p)
                base = os.path.basename(p)
                times.append(base)
                H.append(m["water_depth"])
                QX.append(m["unit_discharge_x"])
                QY.append(m["unit_discharge_y"])
            keep = list(range(len(outs)))
            if name.startswith("dambreak_20k"):
                keep = [len(outs) - 1]  # final map only for the full-size case
            if "_micro_" in name:  # one map, named for t = 0, holding the state at t_max
                meta["times"] = np.array([kw["t_max"]])
            np.savez_compressed(os.path.join(OUT, name + ".npz"), files=np.array([times[k] for k in keep]),
                                h=np.array([H[k] for k in keep]), qx=np.array([QX[k] for k in keep]),
                                qy=np.array([QY[k] for k in keep]), **meta)
            print(name, "maps:", [times[k] for k in keep])
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    if not only:
        main_e(sw)


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--e-only":  # regenerate named _E fixtures only
        sys.path.insert(0, REPO)
        import bench
        main_e(bench.load_pkg(), only=set(sys.argv[2:]))
    elif len(sys.argv) > 2 and sys.argv[1] == "--only":  # regenerate named fixtures only
        main(only=set(sys.argv[2:]))
    else:
        main()
