#!/usr/bin/env python3
"""Benchmark of the SP hot path on one MI355X: forward + adjoint.

Workload (the calibration gradient on the 1M mesh): 5 x 5 km jittered-lattice
Delaunay mesh (707 x 707 lattice -> 999,698 triangles), bed slope 1e-3 + 20
seeded Gaussian bumps, porosity in [0.3, 1], partially dry start (~40 %
wet), Godunov + HLLC, Strickler-law friction of the differentiable forward
with 8 Manning zones (k-means of the centroids), a 60-knot west inflow
hydrograph and 20 depth gauges observed every 30 s over a 90 s window
(~500 time steps).  The observations come from a twin run at seeded "true"
controls; the gradient is taken at the default controls (n = 0.03, the
unperturbed hydrograph).

One "step" of this benchmark is one gradient evaluation: the dry run that
fixes the time-step schedule (with sqrt(N) checkpoints), the forward
recomputation of each segment and the reverse sweep -- i.e. one forward and
one adjoint pass over every time step of the window.  `value` counts
fwd+adj cell-steps: n_cell x time steps x gradient evaluations / wall time.
All inputs are resident in HBM when the timed region starts; each call
copies the 3 x n_cell control arrays in and the gradients out (<0.1 % of
the time).

The same gradient with the MUSCL-EVR _E forward (limited free-surface
slopes, desingularised velocities) and its adjoint is reported under
"fwd_adj_muscl_evr" (or the Godunov one there with --scheme muscl_evr).
The MUSCL-EVR adjoint grows through this window's wet/dry fronts until the
reference's explosion guard (ratio 1e8) would stop the reverse sweep part
way; that line is timed with the guard off so that every time step is
reversed, and the guarded gradient's first kept step is reported beside it.
The native (non-differentiable) MUSCL-EVR forward on the same mesh is timed
as well and reported under "forward".

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(REPO, "inria-internship-project-sw2d-enzyme_amd")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E vendor peak
METRIC = "FP64 cell-steps/s fwd+adj (%HBM roofline); 4D-Var time-to-misfit learned vs plain"
SCHEME_TEXT = {"godunov": "Godunov/HLLC", "muscl_evr": "MUSCL-EVR (limited free-surface slopes)/HLLC"}


def load_pkg():
    spec = importlib.util.spec_from_file_location("sw2d_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    m = importlib.util.module_from_spec(spec)
    sys.modules["sw2d_amd"] = m
    spec.loader.exec_module(m)
    return m


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def pass_bytes(nc, ni, nb, scheme):
    """Compulsory HBM bytes of one time step (inputs read once, outputs
    written once, no intermediates).
      forward  Godunov   80 Nc + 32 Ei + 32 Eb,  MUSCL-EVR 96 Nc + 48 Ei + 32 Eb
      reverse  Godunov  120 Nc + 32 Ei + 32 Eb,  MUSCL-EVR 120 Nc + 48 Ei + 32 Eb
    per cell: state 24 in + 24 out + {zb, A, phi, n} 32 (+ centroid 16, MUSCL);
    reverse per cell: primal state 24 + statics 48 + adjoint 24 in + 24 out;
    per internal edge: 2 int32 cells + {nx, ny, L} 24 (+ midpoint 16, MUSCL);
    per boundary edge: int32 cell + int32 BC + {nx, ny, L} 24."""
    if scheme == "muscl_evr":
        return 96 * nc + 48 * ni + 32 * nb, 120 * nc + 48 * ni + 32 * nb
    return 80 * nc + 32 * ni + 32 * nb, 120 * nc + 32 * ni + 32 * nb


def kernel_bytes(name, nc, ni, nb, n_gauges, scheme="godunov"):
    """Compulsory bytes of one launch of each kernel of the gradient (every
    array element the launch touches, read once / written once; int32
    indices, FP64 values; friction on, head loss off).  <true> variants are
    the MUSCL-EVR ones."""
    muscl = scheme == "muscl_evr"
    table = {
        # cells U, zb, beta, area, CSR ptrs (+ slopes, centroids); incidences,
        # edge cells and geometry (+ midpoints); boundary incidences and geometry
        "k_e_cfl<false>": 56 * nc + 40 * ni + 28 * nb,
        "k_e_cfl<true>": 88 * nc + 56 * ni + 28 * nb,
        "k_e_dt": 8 * ((nc + 255) // 256),
        # fused flux + update: cells U in/out, zb, beta, phi, area, K1, K2, cos,
        # sin, ptrs (+ slopes, centroids); incidences, their neighbours and
        # edge-side factors (Godunov), edge geometry (+ midpoints); boundary
        # incidences, geometry, series; the replay
        # (<true, ., .>) also writes the balanced state + increments and the
        # boundary fluxes.  Each internal edge's flux is computed by both of its
        # cells; the bytes count its inputs once.
        "k_e_update<false, false, false>": 120 * nc + 48 * ni + 32 * nb,
        "k_e_update<true, false, false>": 168 * nc + 48 * ni + 56 * nb,
        # MUSCL-EVR: edges once (states, fluxes and, dry run, swept terms
        # out), the CFL from the swept terms, the update from the stored
        # fluxes (F1 and each side's momentum sum, 40 bytes per edge) and the
        # edge length; the replay's update also writes the balanced state +
        # increments
        "k_e_edges_muscl": 88 * nc + 80 * ni + 48 * nb,
        "k_e_cfl_sw": 56 * nc + 24 * ni + 28 * nb,
        "k_e_update<false, true, true>": 104 * nc + 56 * ni + 52 * nb,
        "k_e_update<true, true, true>": 152 * nc + 56 * ni + 52 * nb,
        # cells U(h), zb, centroid, CSR ptr, slopes out; incidences, edge cells
        # and midpoints
        "k_e_slopes": 52 * nc + 32 * ni,
        # cells: the cell part in (lams, or part for a boundary cell), ptrs
        # (internal and boundary), P in, phi, K1, K2, cos, sin, lams out,
        # gphi/gK1/gK2 read-modify-write; incidences + edge sides (+ the
        # neighbours' slope terms, MUSCL)
        "k_r_gc<true>": 196 * nc + (80 if muscl else 72) * ni,
        "k_r_gc<false>": 84 * nc + (80 if muscl else 72) * ni,
        # cells phi, zb, beta, U, lams; edges L, R, nx, ny, the two sides'
        # factors; sides out; the boundary blocks: boundary cells' state,
        # statics, lams, part, gphi, boundary edges' geometry, F, gbce
        "k_r_edge<false>": 72 * nc + 104 * ni + 160 * nb,
        # + cell slopes and centroids, edge midpoints, slope adjoints out
        "k_r_edge<true>": 104 * nc + 152 * ni + 160 * nb,
        # cells U(h), zb, centroid, ptr, part(h) in/out; incidences, slope
        # adjoints in, neighbour terms out, edge cells and midpoints
        "k_r_slopes": 52 * nc + 80 * ni,
        "k_gauge_sample": 16 * n_gauges,
        "k_cost_gauges": 32 * n_gauges,
        "k_cost_total": 64,
    }
    return table.get(name)


def fwd_adj_bench(sw, args, scheme, explosion_ratio=1e8):
    """Gradients on the calibration workload; explosion_ratio = 0 times the
    whole reverse sweep even where the guard would stop it (the guarded
    gradient's first kept step is reported as well)."""
    from sw2d_amd import assim, calibration
    t0 = time.perf_counter()
    cal = calibration.Calibration(px=args.px, window=args.window, scheme=scheme)
    case = cal.case
    nc, ni, nb = case.n_cell, case.n_int, case.n_bnd
    log(f"{scheme}: calibration mesh {nc} cells, {ni} internal edges, {nb} boundary edges "
        f"(built in {time.perf_counter() - t0:.1f} s)")
    s = sw.SpSolver(case, device=0)
    # twin run at the true controls -> gauge observations
    K_true = cal.K(cal.n_true)
    cal.apply_hydrograph(s, cal.Q_true)
    ng, no = len(cal.gauge_cells), len(cal.obs_t)
    T = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=np.zeros((no, ng)),
                    max_steps=args.max_steps)
    T.record(K1=K_true, K2=K_true)
    obs = T.gauge_series()
    # gradient at the default controls
    cal.apply_hydrograph(s, cal.Q_default)
    K0 = cal.K(cal.n_default)
    guarded = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs,
                          max_steps=args.max_steps)
    guarded.gradient(K1=K0, K2=K0)
    kept = guarded.truncation()[0]
    A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs,
                    max_steps=args.max_steps, explosion_ratio=explosion_ratio)
    for _ in range(args.warmup):
        G = A.gradient(K1=K0, K2=K0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        G = A.gradient(K1=K0, K2=K0)  # returns after a stream synchronisation
    elapsed = time.perf_counter() - t0
    n_steps = G["n_steps"]
    rec = A.get_record()
    gn, gQ = cal.grad_controls(G, cal.n_default, rec["t"][:-1])
    if A.truncation()[0] != 0:
        raise RuntimeError("the timed gradients did not reverse every time step")
    log(f"{scheme}: guarded gradient keeps steps {kept}..{n_steps - 1} (explosion guard, ratio 1e8)")
    log(f"{scheme}: {args.steps} gradients x {n_steps} time steps in {elapsed:.3f} s; J = {G['J']:.6e}, "
        f"|dJ/dn| = {np.linalg.norm(gn):.3e}, |dJ/dQ| = {np.linalg.norm(gQ):.3e}")
    prof = None
    if not args.no_profile:
        log(f"{scheme}: per-launch timing of one gradient (HIP events)")
        prof = A.profile(K1=K0, K2=K0)
    out = dict(scheme=scheme, guard_first_kept_step=kept, explosion_ratio=explosion_ratio,
               cal=cal, case=case, solver=s, assim=A, n_steps=n_steps, elapsed=elapsed,
               J=G["J"], prof=prof, rec=rec, obs=obs, K0=K0, gn=gn, gQ=gQ,
               stride=int(np.ceil(np.sqrt(args.max_steps))))
    return out


def roofline(res, args):
    scheme = res["scheme"]
    case = res["case"]
    nc, ni, nb = case.n_cell, case.n_int, case.n_bnd
    n_steps = res["n_steps"]
    ng = len(res["cal"].gauge_cells)
    kernels = {}
    total_ms = 0.0
    for name, (avg, cnt) in res["prof"].items():
        total_ms += avg * cnt
        b = kernel_bytes(name, nc, ni, nb, ng, scheme)
        kernels[name] = {"avg_ms": avg, "launches": cnt, "bytes_per_launch": b,
                         "gbs": (b / (avg * 1e-3) / 1e9) if (b and avg > 0) else None}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    d = kernels[dom]
    Bf, Ba = pass_bytes(nc, ni, nb, scheme)
    # achieved = the algorithmic bytes of the dominant kernel's pass (B_f for a
    # forward kernel, B_a for a reverse one: one launch advances one time step
    # of that pass) over its average launch time
    reverse = dom.startswith("k_r_")
    B_pass = Ba if reverse else Bf
    achieved = B_pass / (d["avg_ms"] * 1e-3) / 1e9
    step_ms = total_ms / n_steps  # device time per time step (fwd + recompute + adj)
    step_gbs = (Bf + Ba) / (step_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_fwd_adj_1m.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(scheme, {}).get(dom)
    impl = None
    if d["bytes_per_launch"]:
        impl = {"bytes_per_launch": d["bytes_per_launch"], "achieved": d["gbs"],
                "frac": d["gbs"] / HBM_PEAK_GBS,
                "note": "every array element the launch touches, read once / written once "
                        "(the kernel's own inputs and outputs, intermediates included)"}
    return {
        "bound": "hbm",
        "kernel": dom,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "algorithmic_bytes_per_launch": B_pass,
        "pass": "reverse (B_a)" if reverse else "forward (B_f)",
        "avg_launch_ms": d["avg_ms"],
        "implementation": impl,
        "step": {
            "algorithmic_bytes_fwd_adj": Bf + Ba,
            "device_ms_per_time_step": step_ms,
            "achieved": step_gbs,
            "frac": step_gbs / HBM_PEAK_GBS,
            "note": "algorithmic fwd+adj bytes of one time step over the device time of all "
                    "kernels of one time step (dry run + recomputation + reverse)",
        },
    }, kernels


def forward_bench(sw, args):
    case, info = sw.scale_case(None, px=args.px, py=args.px, L=5000.0, seed=1234, t_max=3600.0,
                               scheme="muscl_evr")
    s = sw.SpSolver(case, device=0)
    s.run(max_steps=64)
    t0 = time.perf_counter()
    n = s.run(max_steps=args.fwd_steps)
    el = time.perf_counter() - t0
    prof = s.profile(64)
    Bf, _ = pass_bytes(case.n_cell, case.n_int, case.n_bnd, "muscl_evr")
    step_ms = sum(prof.values())
    s.close()
    log(f"native MUSCL forward: {n} steps in {el:.3f} s")
    cpu = None
    if not args.no_cpu:
        log("native forward: CPU baselines (OpenMP and serial oracle)")
        cpu = cpu_forward_baseline(case, args)
    traffic = None  # HBM bytes of one step from the copy-calibrated counters
    tpath = os.path.join(REPO, "profiles", "traffic_fwd_adj_1m.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tk = json.load(f).get("forward_muscl_evr", {})
        if all(k in tk for k in ("k_slopes", "k_edges<1>", "k_cfl", "k_update<false>")):
            traffic = sum(tk[k] for k in ("k_slopes", "k_edges<1>", "k_cfl", "k_dt", "k_update<false>") if k in tk)
    return {"workload": "SP forward at scale, MUSCL-EVR + Manning 0.03, native semantics",
            "cpu_baseline": cpu,
            "value": case.n_cell * n / el, "unit": "cell-steps/s", "steps": n,
            "ms_per_step": el * 1e3 / n, "kernels_ms": prof,
            "roofline_step": {"algorithmic_bytes": Bf, "device_ms": step_ms,
                              "achieved": Bf / (step_ms * 1e-3) / 1e9,
                              "frac": Bf / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "traffic": traffic,
                              "note": "algorithmic bytes of one step over the summed per-kernel "
                                      "device time of one step; traffic: the kernels' HBM bytes "
                                      "per step (profiles/counters_forward_1m.md)"}}


def host_info():
    """The host the CPU baselines ran on: logical CPUs available to this
    process and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return {"nproc": n, "cpu_model": model}


def _oracle():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    return O


def cpu_baseline(res, args):
    """The oracle's _E dry run + reverse sweep (all states kept, no
    recomputation) on a prefix of the same gradient: the OpenMP build on
    every host thread it is given (OMP_NUM_THREADS), and the serial build
    (1 thread, the reference is single-threaded) on a shorter prefix."""
    O = _oracle()
    case, cal, K0 = res["case"], res["cal"], res["K0"]

    def leg(math, m):
        E = O.OracleE(case, math=math)
        t_m = float(res["rec"]["t"][m])
        tc = time.perf_counter()
        rec = E.record([t_m], K1=K0, K2=K0, max_steps=m)
        g = E.gradient(rec, K1=K0, K2=K0, gauges=cal.gauge_cells, obs=res["obs"][:1], n_obs=1)
        tc = time.perf_counter() - tc
        n = len(rec["dt"])
        return case.n_cell * n / tc, n, tc, g["J"]

    threads = O.lib("omp").orc_threads()
    v, n, tc, J = leg("omp", args.cpu_steps_omp)
    v1, n1, tc1, J1 = leg("glibc", args.cpu_steps)
    return {
        "value": v,
        "unit": "fwd+adj cell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} time steps of the same {case.n_cell}-cell gradient (C oracle, OpenMP "
                  f"build: _E dry run + reverse sweep, all states kept, no recomputation), "
                  f"{threads} threads, {tc:.1f} s; J = {J:.6e}",
        "host": host_info(),
        "single_thread": {"value": v1, "cores": 1, "sample": f"first {n1} time steps, serial build, "
                                                             f"{tc1:.1f} s; J = {J1:.6e}"},
    }


def cpu_forward_baseline(case, args):
    """The oracle's native forward (the GPU line's workload) on a bounded
    number of steps: OpenMP build on every host thread, serial build on 1."""
    O = _oracle()

    def leg(math, m):
        o = O.OracleSolver(case, math=math)
        tc = time.perf_counter()
        o.step(m)
        return case.n_cell * m / (time.perf_counter() - tc), time.perf_counter() - tc

    threads = O.lib("omp").orc_threads()
    v, tc = leg("omp", args.cpu_fwd_steps_omp)
    v1, tc1 = leg("glibc", args.cpu_fwd_steps)
    return {"value": v, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {args.cpu_fwd_steps_omp} steps of the same {case.n_cell}-cell native "
                      f"MUSCL-EVR forward (C oracle, OpenMP build), {threads} threads, {tc:.1f} s",
            "host": host_info(),
            "single_thread": {"value": v1, "cores": 1,
                              "sample": f"first {args.cpu_fwd_steps} steps, serial build, {tc1:.1f} s"}}


def time_to_misfit_bench(sw, args):
    """The learned-vs-plain 4D-Var experiment: ensemble training set -> MLP
    -> 4D-Var on the px-lattice mesh from the network's prediction and from
    the default controls, both to the same gauge-misfit tolerance."""
    from sw2d_amd import calibration, warmstart
    lo, hi = (float(v) for v in args.train_cells.split(","))
    if args.ttm_stage in ("all", "train"):
        ts = warmstart.training_set(args.train_runs, launch_runs=args.launch_runs,
                                    n_meshes=args.train_meshes, cells=(lo, hi),
                                    window=args.ttm_window, log=log)
        ws = warmstart.WarmStart(ts["Q_default"])
        t0 = time.perf_counter()
        hist = ws.fit(ts["gauges"], ts["h0"], ts["n"], ts["Q"], epochs=args.epochs, log=log)
        stats = {"ens_runs_per_launch": ts["runs_per_launch"], "ens_launches": ts["launches"],
                 "ens_block": ts["block"], "ens_device_s": ts["device_ms"] / 1e3,
                 "ens_cell_steps_per_s": ts["cell_steps"] / (ts["device_ms"] * 1e-3),
                 "ens_wall_s": ts["wall_s"], "fit_s": time.perf_counter() - t0,
                 "train_loss": hist[-1][0], "validation_loss": hist[-1][1],
                 "train_runs": args.train_runs, "train_cells_lo": lo, "train_cells_hi": hi}
        if args.ttm_cache:
            ws.save(args.ttm_cache, **stats)
        if args.ttm_stage == "train":
            return {"leg": "learned warm start: training", **stats}
    else:
        ws, d = warmstart.WarmStart.load(args.ttm_cache)
        stats = {k: float(d[k]) for k in d if k.startswith(("ens_", "fit_", "train", "valid"))}
    cal = calibration.Calibration(px=args.px, window=args.ttm_window, level_quantile=0.9)
    s = sw.SpSolver(cal.case, device=0)
    t0 = time.perf_counter()
    obs = cal.observe(s, cal.n_true, cal.Q_true, max_steps=args.ttm_max_steps)
    log(f"observed the truth on {cal.case.n_cell} cells in {time.perf_counter() - t0:.1f} s")
    h0g = cal.case.a["h0"][cal.gauge_cells]
    t0 = time.perf_counter()
    n_l, Q_l = ws.predict(obs, h0g)
    pred_s = time.perf_counter() - t0
    n_obs, n_g = obs.shape
    J_tol = args.misfit_rms ** 2 * n_obs * n_g
    log(f"learned start: max |n/n_true - 1| = {np.max(np.abs(n_l / cal.n_true - 1)):.3f} "
        f"(default {np.max(np.abs(cal.n_default / cal.n_true - 1)):.3f}); J_tol = {J_tol:.3e}")
    starts = {"learned": (n_l, Q_l, pred_s), "plain": (cal.n_default, cal.Q_default, 0.0)}
    res = warmstart.time_to_misfit(cal, s, obs, starts, J_tol, max_iter=args.ttm_max_iter,
                                   time_limit=args.ttm_time_limit,
                                   max_steps=args.ttm_max_steps, n_tiles=args.ttm_tiles,
                                   adj_cut=args.adj_cut, log=log)
    s.close()
    tl, tp = res["learned"]["seconds"], res["plain"]["seconds"]
    return {
        "metric": METRIC,
        "leg": "4D-Var time-to-misfit, learned vs plain",
        "value": (tp / tl) if (tl and tp) else None,
        "unit": "plain / learned time to the misfit tolerance",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "f64 (4D-Var), bf16 MFMA / fp32 accumulate (network)",
        "data": "synthetic: seeded meshes, prior-drawn controls, twin observations",
        "config": {
            "workload": f"4D-Var on the {cal.case.n_cell}-cell mesh, {args.ttm_window:.0f} s window "
                        f"in {args.ttm_tiles} tile(s), adjoint thin-film cut {args.adj_cut} m, "
                        f"68 controls (8 Manning zones + 60-knot "
                        f"hydrograph), 20 gauges every 30 s; tolerance: RMS gauge misfit "
                        f"{args.misfit_rms} m (J <= {J_tol:.4g})",
            "network": [n_obs * n_g, *warmstart.HIDDEN, len(n_l) + len(Q_l)],
        },
        "training": stats,
        "learned": {k: v for k, v in res["learned"].items() if k != "trace"},
        "plain": {k: v for k, v in res["plain"].items() if k != "trace"},
        "traces": {k: v["trace"] for k, v in res.items()},
    }


def ttm_reduced(sw, args):
    """The metric's second half at a reduced size, measured in this run: a
    training set of args.ttm_small_runs ensemble runs (16 meshes of 2e4-3e4
    cells, 3600 s each: 20 gauges x 120 samples), the MLP 2400-512-512-512-68
    fitted to it, then FP64 4D-Var on the px-122 mesh (29,768 cells, 3600 s
    window) from the network's prediction and from the default controls,
    both to an RMS gauge misfit of 5 mm.  The learned leg's time includes the
    network's inference; training is reported, not charged to it (the
    network serves every later calibration of the problem)."""
    from sw2d_amd import calibration, warmstart
    t0 = time.perf_counter()
    ts = warmstart.training_set(args.ttm_small_runs, launch_runs=1024, n_meshes=16,
                                cells=(2.0e4, 3.0e4), window=3600.0, log=log)
    ws = warmstart.WarmStart(ts["Q_default"])
    t1 = time.perf_counter()
    hist = ws.fit(ts["gauges"], ts["h0"], ts["n"], ts["Q"], epochs=60)
    fit_s = time.perf_counter() - t1
    cal = calibration.Calibration(px=122, window=3600.0, level_quantile=0.9)
    s = sw.SpSolver(cal.case, device=0)
    obs = cal.observe(s, cal.n_true, cal.Q_true, max_steps=60000)
    t2 = time.perf_counter()
    n_l, Q_l = ws.predict(obs, cal.case.a["h0"][cal.gauge_cells])
    pred_s = time.perf_counter() - t2
    n_obs, n_g = obs.shape
    J_tol = 0.005 ** 2 * n_obs * n_g
    starts = {"learned": (n_l, Q_l, pred_s), "plain": (cal.n_default, cal.Q_default, 0.0)}
    res = warmstart.time_to_misfit(cal, s, obs, starts, J_tol, max_iter=60, time_limit=40.0,
                                   max_steps=60000, adj_cut=1e-3, log=log)
    s.close()
    tl, tp = res["learned"]["seconds"], res["plain"]["seconds"]
    return {
        "measured_in_this_run": True,
        "value": (tp / tl) if (tl and tp) else None,
        "unit": "plain / learned seconds to the misfit tolerance",
        "learned_s": tl, "plain_s": tp,
        "gradients": {"learned": res["learned"]["grads"], "plain": res["plain"]["grads"]},
        "J0": {"learned": res["learned"]["J0"], "plain": res["plain"]["J0"]},
        "workload": f"4D-Var on the {cal.case.n_cell}-cell mesh, 3600 s window, 68 controls "
                    f"(8 Manning zones + 60-knot hydrograph), 20 gauges every 30 s; tolerance "
                    f"RMS gauge misfit 5 mm (J <= {J_tol:.4g}); adjoint thin-film cut 1 mm; "
                    f"40 s limit per start",
        "network": list(ws.net.dims),
        "training": {"runs": int(len(ts["n"])), "mesh_cells": [20000, 30000], "window_s": 3600.0,
                     "ensemble_device_s": ts["device_ms"] / 1e3, "ensemble_wall_s": ts["wall_s"],
                     "fit_s": fit_s, "train_loss": hist[-1][0], "validation_loss": hist[-1][1]},
        "leg_wall_s": time.perf_counter() - t0,
        "full_size": "python bench.py --time-to-misfit (1M mesh; last record in "
                     "profiles/time_to_misfit_1m.md)",
    }


def ensemble_bench(sw, args):
    """The batched-ensemble leg: 256 runs of the calibration problem on
    meshes of 20k-250k cells (log-uniform), 1200 s each, one launch, the
    workgroup size picked by the autotuner."""
    from sw2d_amd import ensemble
    t0 = time.perf_counter()
    P = ensemble.EnsembleProblem(n_runs=args.ens_runs, seed=7, window=1200.0)
    log(f"ensemble: {len(P.cals)} runs, {P.n_cells} cells ({P.ens.device_bytes() / 1e9:.2f} GB), "
        f"built in {time.perf_counter() - t0:.1f} s")
    n, Q = P.sample(1)
    P.set_controls(n, Q)
    tuner = ensemble.Autotuner(probe_steps=30)
    block = tuner.pick(P.ens, P.obs_t)
    log(f"autotune (ms per step of the slowest run): {tuner.last} -> {block}")
    t0 = time.perf_counter()
    g, steps, ms = P.ens.run(P.obs_t, block=block)
    wall = time.perf_counter() - t0
    wgs, cmax = P.ens.last_launch()
    cells = np.array([c.case.n_cell for c in P.cals])
    cs = int(np.sum(cells * steps))
    log(f"ensemble run: {ms / 1e3:.2f} s device ({wall:.2f} s wall), {wgs} workgroups, clusters up "
        f"to {cmax}; {cs / (ms * 1e-3):.3e} cell-steps/s")
    return {"metric": METRIC, "leg": "batched ensemble (_E forward, gauge series)",
            "value": cs / (ms * 1e-3), "unit": "cell-steps/s", "higher_is_better": True,
            "n_gpus": 1, "dtype": "f64",
            "data": "synthetic: seeded meshes, prior-drawn controls",
            "config": {"workload": f"{len(P.cals)} runs of the calibration problem, meshes "
                                   f"{int(cells.min())}-{int(cells.max())} cells, 1200 s each, "
                                   "20 gauges x 40 observations per run",
                       "total_cells": int(cells.sum()), "cell_steps": cs,
                       "steps_min_max": [int(steps.min()), int(steps.max())]},
            "device_s": ms / 1e3, "wall_s": wall, "block": block, "workgroups": wgs,
            "max_cluster": cmax, "autotune_ms_per_step": {str(k): v for k, v in tuner.last.items()}}


def mlp_bench(sw, args):
    """The warm start's network 2400-512-512-512-68 on synthetic data:
    Adam steps of 256 rows on a 15,360-row training set resident on the
    device (an epoch of 60 steps per call, one wait each), and the forward
    of 16,384 rows (host arrays in and out)."""
    from sw2d_amd import mlp
    dims = [2400, 512, 512, 512, 68]
    rng = np.random.default_rng(3)
    n, batch = 15360, 256
    x = rng.standard_normal((n, dims[0])).astype(np.float32)
    t = rng.standard_normal((n, dims[-1])).astype(np.float32)
    net = mlp.MLP(dims, max_batch=16384, seed=1)
    t0 = time.perf_counter()
    net.set_data(x, t)
    t_data = time.perf_counter() - t0
    net.train_resident(epochs=1, batch=batch, lr=1e-4, seed=0)
    reps = 5
    t0 = time.perf_counter()
    net.train_resident(epochs=reps, batch=batch, lr=1e-4, seed=1)
    t_ep = (time.perf_counter() - t0) / reps
    steps = n // batch
    flop_row = 2 * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
    # forward 1x, backward 2x (weight and input gradients; the first layer's input gradient is not formed)
    flop_step = batch * (3 * flop_row - 2 * dims[0] * dims[1])
    xf = rng.standard_normal((16384, dims[0])).astype(np.float32)
    net.forward(xf)
    t0 = time.perf_counter()
    for _ in range(3):
        net.forward(xf)
    t_fwd = (time.perf_counter() - t0) / 3
    net.close()
    ms_step = t_ep * 1e3 / steps
    log(f"mlp: {ms_step * 1e3:.1f} us per training step of {batch} rows "
        f"({flop_step / (ms_step * 1e-3) / 1e12:.1f} TFLOP/s), 16384-row forward {t_fwd * 1e3:.2f} ms")
    return {"leg": "MLP of the learned warm start (BF16 MFMA, FP32 accumulation)",
            "network": dims, "batch": batch, "training_rows": n, "steps_per_epoch": steps,
            "ms_per_training_step": ms_step, "training_tflops": flop_step / (ms_step * 1e-3) / 1e12,
            "set_data_s": t_data, "forward_16384_rows_ms": t_fwd * 1e3,
            "note": "host-timed, one wait per epoch; the forward includes the host-to-device copy "
                    "of its 157 MB of inputs; per-kernel GEMM times come from a kernel trace"}


def copy_calibration(args):
    """A plain device copy of known size (torch's elementwise copy kernel,
    2 x 2^28 doubles read and written once) for scaling the FETCH_SIZE /
    WRITE_SIZE counters of a profiled run to bytes."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    return {"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed gradient evaluations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--px", type=int, default=707, help="lattice size (707 -> ~1e6 cells)")
    ap.add_argument("--window", type=float, default=90.0, help="assimilation window (s)")
    ap.add_argument("--max-steps", type=int, default=2000)
    ap.add_argument("--fwd-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=40,
                    help="time steps of the serial CPU fwd+adj baseline")
    ap.add_argument("--cpu-steps-omp", type=int, default=400,
                    help="time steps of the OpenMP CPU fwd+adj baseline")
    ap.add_argument("--cpu-fwd-steps", type=int, default=60,
                    help="steps of the serial CPU native-forward baseline")
    ap.add_argument("--cpu-fwd-steps-omp", type=int, default=1000,
                    help="steps of the OpenMP CPU native-forward baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forward", action="store_true")
    ap.add_argument("--scheme", choices=("muscl_evr", "godunov"), default="godunov",
                    help="reconstruction of the headline fwd+adj line")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the per-launch event timing and the roofline (for counter runs)")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the fwd+adj line of the other reconstruction")
    ap.add_argument("--no-ttm", action="store_true",
                    help="skip the reduced time-to-misfit leg of the default run")
    ap.add_argument("--ttm-small-runs", type=int, default=1024,
                    help="training-set size of the reduced time-to-misfit leg")
    ap.add_argument("--ensemble", action="store_true",
                    help="run the batched-ensemble leg instead")
    ap.add_argument("--ens-runs", type=int, default=256)
    ap.add_argument("--mlp", action="store_true",
                    help="only the MLP leg (training steps and a large forward)")
    ap.add_argument("--forward-only", action="store_true",
                    help="only the native MUSCL-EVR forward line (for kernel traces and counters)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain copy kernel of known size (counter scaling)")
    ap.add_argument("--time-to-misfit", action="store_true",
                    help="run the learned-vs-plain 4D-Var experiment instead")
    ap.add_argument("--train-runs", type=int, default=16384)
    ap.add_argument("--launch-runs", type=int, default=1024)
    ap.add_argument("--train-meshes", type=int, default=64)
    ap.add_argument("--train-cells", default="2e4,1e5")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--ttm-window", type=float, default=1200.0)
    ap.add_argument("--ttm-max-steps", type=int, default=30000)
    ap.add_argument("--ttm-max-iter", type=int, default=200)
    ap.add_argument("--ttm-time-limit", type=float, default=420.0)
    ap.add_argument("--misfit-rms", type=float, default=0.01)
    ap.add_argument("--ttm-tiles", type=int, default=1)
    ap.add_argument("--adj-cut", type=float, default=1e-3,
                    help="thin-film cut of the 4D-Var adjoint (m; 0 = exact adjoint)")
    ap.add_argument("--ttm-stage", choices=("all", "train", "fourdvar"), default="all")
    ap.add_argument("--ttm-cache", default=None,
                    help="npz of the trained network (written by 'train', read by 'fourdvar')")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args)))
        return
    sw = load_pkg()
    if args.time_to_misfit:
        print(json.dumps(time_to_misfit_bench(sw, args)))
        return
    if args.ensemble:
        print(json.dumps(ensemble_bench(sw, args)))
        return
    if args.forward_only:
        print(json.dumps(forward_bench(sw, args)))
        return
    if args.mlp:
        print(json.dumps(mlp_bench(sw, args)))
        return
    ratio = lambda sc: 0.0 if sc == "muscl_evr" else 1e8
    res = fwd_adj_bench(sw, args, args.scheme, ratio(args.scheme))
    if args.no_profile:  # timing only (runs under a counter profiler)
        print(json.dumps({"scheme": args.scheme, "value": res["case"].n_cell * res["n_steps"] * args.steps
                          / res["elapsed"], "unit": "fwd+adj cell-steps/s"}))
        return
    roof, kernels = roofline(res, args)
    case = res["case"]
    nc = case.n_cell
    n_steps = res["n_steps"]
    el = res["elapsed"]
    line = {
        "metric": METRIC,
        "value": nc * n_steps * args.steps / el,
        "unit": "fwd+adj cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded 1M-cell jittered-lattice Delaunay mesh built in memory; "
                "gauge observations from a twin run at seeded true controls",
        "config": {
            "workload": "adjoint gradient on the 1M mesh: 8 Manning zones + 60-knot inflow "
                        "hydrograph (68 controls), 20 gauges every 30 s, 90 s window, "
                        "sqrt(N) checkpointing, " + SCHEME_TEXT[args.scheme] + " _E forward",
            "scheme": args.scheme,
            "n_cell": nc, "n_int": case.n_int, "n_bnd": case.n_bnd,
            "time_steps_per_gradient": n_steps,
            "checkpoint_stride": res["stride"],
            "explosion_guard_ratio": res["explosion_ratio"],
            "guarded_gradient_first_kept_step": res["guard_first_kept_step"],
            "cost_J": res["J"],
        },
        "roofline": roof,
        "kernels": kernels,
    }
    res["solver"].close()
    if not args.no_other:  # the other reconstruction on the same workload
        other = "godunov" if args.scheme == "muscl_evr" else "muscl_evr"
        r2 = fwd_adj_bench(sw, args, other, ratio(other))
        roof2, k2 = roofline(r2, args)
        line["fwd_adj_" + other] = {
            "value": r2["case"].n_cell * r2["n_steps"] * args.steps / r2["elapsed"],
            "unit": "fwd+adj cell-steps/s", "ms_per_step": r2["elapsed"] * 1e3 / args.steps,
            "time_steps_per_gradient": r2["n_steps"], "cost_J": r2["J"],
            "explosion_guard_ratio": r2["explosion_ratio"],
            "guarded_gradient_first_kept_step": r2["guard_first_kept_step"],
            "workload": "the same, " + SCHEME_TEXT[other] + " _E forward",
            "roofline": roof2, "kernels": k2}
        r2["solver"].close()
    if not args.no_ttm:
        line["time_to_misfit"] = ttm_reduced(sw, args)
    if not args.no_forward:
        line["forward"] = forward_bench(sw, args)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(res, args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
