"""Batched ensemble (sw2dg_ens_*) on the GPU: every run of one persistent
launch over uneven meshes reproduces the single-run dry run of the 4D-Var
layer (sw2dg_assim_record, itself checked against the reference's _E
physics) bit for bit -- gauge series, step count and final state -- for
every workgroup size the autotuner may pick, with runs split over clusters
of workgroups (cooperative launch) and with one workgroup per run; and the
batched-ensemble configuration at its stated size (256 runs, 20k-250k
cells, 1200 s each, one launch)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem(sw):
    from sw2d_amd import ensemble
    P = ensemble.EnsembleProblem(n_runs=4, seed=3, window=120.0, sizes=[40, 60, 50, 80])
    return P


def test_gpu_ensemble_runs_match_single_runs_bitwise(sw, problem):
    from sw2d_amd import assim
    P = problem
    print("[gpu] ensemble", [c.case.n_cell for c in P.cals], flush=True)
    n, Q = P.sample(11)
    P.set_controls(n, Q)
    outs = {}
    for block in (256, 512, 1024):
        outs[block] = P.ens.run(P.obs_t, block=block)
        wgs, cmax = P.ens.last_launch()
        print(f"[gpu] block {block}: {wgs} workgroups, clusters up to {cmax}; steps {outs[block][1]}, "
              f"{outs[block][2]:.2f} ms", flush=True)
        if block == 256:
            assert cmax > 1  # the runs were split over clusters
    P.ens.set_cluster_limit(1)
    outs["single"] = P.ens.run(P.obs_t, block=256)
    assert P.ens.last_launch() == (len(P.cals), 1)
    P.ens.set_cluster_limit(64)
    g, steps, _ = outs[256]
    for block in (512, 1024, "single"):
        np.testing.assert_array_equal(outs[block][0], g)
        np.testing.assert_array_equal(outs[block][1], steps)
    for r, c in enumerate(P.cals):
        s = sw.SpSolver(c.case)
        c.apply_hydrograph(s, Q[r])
        K = c.K(n[r])
        A = assim.Assim(s, c.obs_t, gauge_cells=c.gauge_cells,
                        gauge_h=np.zeros((len(c.obs_t), len(c.gauge_cells))))
        assert A.record(K1=K, K2=K) == steps[r]
        np.testing.assert_array_equal(A.gauge_series(), g[r])
        h, qx, qy = A.final_state()
        he, qxe, qye = P.ens.state(r)
        np.testing.assert_array_equal(he, h)
        np.testing.assert_array_equal(qxe, qx)
        np.testing.assert_array_equal(qye, qy)
        s.close()


def test_gpu_ensemble_autotuner_picks_a_candidate(sw, problem, tmp_path):
    from sw2d_amd import ensemble
    P = problem
    tuner = ensemble.Autotuner(path=str(tmp_path / "tune.json"), probe_steps=10)
    b = tuner.pick(P.ens, P.obs_t)
    assert b in ensemble.BLOCKS and set(tuner.last) == set(ensemble.BLOCKS)
    # remembered per ensemble shape
    again = ensemble.Autotuner(path=str(tmp_path / "tune.json"))
    assert again.pick(P.ens, P.obs_t) == b


def test_gpu_ensemble_full_size(sw):
    """256 runs on meshes of 20k-250k cells (log-uniform), each with its own
    prior-drawn Manning zones and hydrograph, 1200 s, one launch: every run
    finishes with finite, non-negative gauge series of 20 gauges x 40
    observations, and the smallest, the largest and two seeded members equal
    their single-run dry runs bit for bit (series, step count, final state)."""
    from sw2d_amd import assim, ensemble
    t0 = time.perf_counter()
    P = ensemble.EnsembleProblem(n_runs=256, seed=7, window=1200.0)
    cells = np.array([c.case.n_cell for c in P.cals])
    print(f"[gpu] ensemble 256 runs, {cells.sum()} cells ({cells.min()}-{cells.max()}), "
          f"built in {time.perf_counter() - t0:.1f} s", flush=True)
    assert len(cells) == 256 and cells.min() >= 19000 and cells.max() <= 252000
    n, Q = P.sample(1)
    P.set_controls(n, Q)
    g, steps, ms = P.ens.run(P.obs_t, block=256)
    wgs, cmax = P.ens.last_launch()
    print(f"[gpu] one launch: {ms / 1e3:.1f} s device, {wgs} workgroups, clusters up to {cmax}, "
          f"steps {steps.min()}-{steps.max()}", flush=True)
    assert g.shape == (256, 40, 20)
    assert np.all(np.isfinite(g)) and np.all(g >= 0.0)
    assert np.all(steps > 0)
    rng = np.random.default_rng(5)
    members = [int(np.argmin(cells)), int(np.argmax(cells))] + [int(r) for r in rng.choice(256, 2, replace=False)]
    for r in members:
        c = P.cals[r]
        s = sw.SpSolver(c.case)
        c.apply_hydrograph(s, Q[r])
        K = c.K(n[r])
        A = assim.Assim(s, c.obs_t, gauge_cells=c.gauge_cells,
                        gauge_h=np.zeros((len(c.obs_t), len(c.gauge_cells))))
        assert A.record(K1=K, K2=K) == steps[r]
        np.testing.assert_array_equal(A.gauge_series(), g[r])
        for a, b in zip(A.final_state(), P.ens.state(r)):
            np.testing.assert_array_equal(b, a)
        print(f"[gpu] run {r} ({cells[r]} cells, {steps[r]} steps): bitwise the single-run dry run",
              flush=True)
        s.close()
    P.ens.close()
