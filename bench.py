#!/usr/bin/env python3
"""Benchmark of the SP hot path on one MI355X.

Workload: SP forward at scale -- 5 x 5 km jittered-lattice Delaunay mesh
(707 x 707 lattice -> 999,698 triangles, 1,498,133 internal edges), bed slope
1e-3 + 20 seeded Gaussian bumps, porosity in [0.3, 1], partially dry start
(~40 % wet), triangular west inflow, Manning 0.03, MUSCL-EVR, FP64.  One
"step" is one time step of the whole mesh (swSp::execute loop body).  The
inputs are built in memory, uploaded once, and are resident in HBM when the
timed region starts; the timed region runs K steps through the HIP-graph
path and ends with a stream synchronisation.

Prints one JSON line (see the keys below).  --gpus N > 1 exits: this build
runs on one GPU.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(REPO, "inria-internship-project-sw2d-enzyme_amd")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E vendor peak


def load_pkg():
    spec = importlib.util.spec_from_file_location("sw2d_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    m = importlib.util.module_from_spec(spec)
    sys.modules["sw2d_amd"] = m
    spec.loader.exec_module(m)
    return m


def algorithmic_bytes(n_cell, n_int, n_bnd, scheme):
    """Compulsory HBM bytes of one forward step (inputs read once, outputs
    written once, no intermediates):
      Godunov   80 Nc + 32 Ei + 32 Eb
      MUSCL-EVR 96 Nc + 48 Ei + 32 Eb
    per cell: state in 24 + state out 24 + {zb, A, phi, n} 32 (+ centroid 16);
    per internal edge: 2 int32 cells + {nx, ny, L} 24 (+ midpoint 16);
    per boundary edge: int32 cell + int32 BC + {nx, ny, L} 24."""
    if scheme == "muscl_evr":
        return 96 * n_cell + 48 * n_int + 32 * n_bnd
    return 80 * n_cell + 32 * n_int + 32 * n_bnd


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--px", type=int, default=707, help="lattice size (707 -> ~1e6 cells)")
    ap.add_argument("--profile-steps", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=60, help="oracle sample (steps of the same mesh)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    sw = load_pkg()
    t_gen = time.perf_counter()
    case, info = sw.scale_case(None, px=args.px, py=args.px, L=5000.0, seed=1234, t_max=3600.0,
                               scheme="muscl_evr")
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] mesh {info['n_cell']} cells, {info['n_int']} internal edges "
          f"(built in {t_gen:.1f} s)", file=sys.stderr, flush=True)
    s = sw.SpSolver(case, device=0)
    s.run(max_steps=args.warmup)
    c0 = s.clock()
    t0 = time.perf_counter()
    n = s.run(max_steps=args.steps)  # ends with hipStreamSynchronize
    elapsed = time.perf_counter() - t0
    c1 = s.clock()
    print(f"[bench] {n} steps in {elapsed:.3f} s, t {c0['t']:.2f} -> {c1['t']:.2f}",
          file=sys.stderr, flush=True)
    nc, ni, nb = case.n_cell, case.n_int, case.n_bnd
    value = nc * n / elapsed
    # per-kernel device durations, HIP events on the solver's stream
    prof = s.profile(args.profile_steps)
    step_ms_events = sum(prof.values())
    B = algorithmic_bytes(nc, ni, nb, "muscl_evr")
    achieved = B / (step_ms_events * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_sp_step_1m.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_step")
    dominant = max(prof, key=prof.get)
    line = {
        "metric": "FP64 cell-steps/s fwd+adj (%HBM roofline); 4D-Var time-to-misfit learned vs plain",
        "value": value,
        "unit": "cell-steps/s (forward)",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(n, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded 1M-cell jittered-lattice Delaunay mesh, built in memory",
        "config": {
            "workload": "SP forward at scale (5x5 km, 999,698-cell Delaunay, MUSCL-EVR, Manning 0.03, "
                        "wet/dry start, west inflow); forward step only",
            "n_cell": nc, "n_int": ni, "n_bnd": nb, "scheme": "muscl_evr",
            "topo": "discontinuous", "sim_time_s": [c0["t"], c1["t"]],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "SP step = " + " + ".join(prof),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_step": B,
            "step_ms_events": step_ms_events,
        },
        "kernels_ms": prof,
        "dominant_kernel": dominant,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        o = O.OracleSolver(case)
        tc = time.perf_counter()
        done = o.step(args.cpu_steps)
        tc = time.perf_counter() - tc
        line["cpu_baseline"] = {
            "value": nc * done / tc,
            "unit": "cell-steps/s (forward)",
            "cores": 1,
            "kind": "port",
            "sample": f"first {done} steps of the same {nc}-cell workload, C oracle "
                      f"(bit-identical to the compiled reference on the golden cases), 1 thread, {tc:.1f} s",
        }
    s.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
