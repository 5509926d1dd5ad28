"""Batched ensemble (sw2dg_ens_*) on the GPU: every run of one persistent
launch over uneven meshes reproduces the single-run dry run of the 4D-Var
layer (sw2dg_assim_record, itself checked against the reference's _E
physics) bit for bit -- gauge series, step count and final state -- for
every workgroup size the autotuner may pick, with runs split over clusters
of workgroups (cooperative launch) and with one workgroup per run."""
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
