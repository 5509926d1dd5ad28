"""The oracle's OpenMP build (liboracle_omp.so: the benchmark's all-core CPU
baseline) against the serial one and the reference: the native forward
reproduces the reference's golden maps bit for bit, the _E dry run is the
serial dry run bit for bit (schedule and every state), and the _E gradient,
whose reverse sweep sums each edge's terms before adding them to the cell,
agrees with the serial sweep to 1e-12 of each control kind's largest entry
(the cost, computed by the forward, is bit-identical)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import e_golden_names, golden_names, load_golden, map_time, regen_case

NATIVE = [n for n in golden_names() if not n.endswith("_error")]


@pytest.fixture(scope="module", autouse=True)
def threads():
    os.environ.setdefault("OMP_NUM_THREADS", "4")


def test_omp_build_runs_threads(oracle):
    assert oracle.lib("omp").orc_threads() >= 2
    assert oracle.lib("glibc").orc_threads() == 1


@pytest.mark.parametrize("name", NATIVE)
def test_omp_native_matches_reference_maps(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    o = oracle.OracleSolver(sw.Case(str(tmp_path)), math="omp")
    for k, f in enumerate(g["files"]):
        t = map_time(g, k)
        if t > 0:
            o.run(t_end=t)
        h, qx, qy = o.state()
        np.testing.assert_array_equal(h, g["h"][k], err_msg=f"{name} h at {f}")
        np.testing.assert_array_equal(qx, g["qx"][k], err_msg=f"{name} qx at {f}")
        np.testing.assert_array_equal(qy, g["qy"][k], err_msg=f"{name} qy at {f}")


def _gauges(case, n_obs):
    h0 = case.a["h0"]
    cells = np.nonzero(h0 > 0.1)[0][::5][:8].astype(np.int64)
    obs = np.array([[h0[c] * (1.0 + 0.1 * (k + 1)) for c in cells] for k in range(n_obs)])
    return cells, obs


@pytest.mark.parametrize("name", e_golden_names())
def test_omp_e_dry_run_and_gradient(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    obs_t = list(g["obs_t"])
    S, P = oracle.OracleE(case), oracle.OracleE(case, math="omp")
    rs = S.record(obs_t, states=True)
    rp = P.record(obs_t, states=True)
    np.testing.assert_array_equal(rp["dt"], rs["dt"])
    np.testing.assert_array_equal(rp["trig"], rs["trig"])
    np.testing.assert_array_equal(rp["states"], rs["states"])
    cells, obs = _gauges(case, len(obs_t))
    rr = {"dt": rs["dt"], "bc": rs["bc"], "trig": rs["trig"]}
    Gs = S.gradient(rr, gauges=cells, obs=obs, normalize=True, n_obs=len(obs_t))
    Gp = P.gradient(rr, gauges=cells, obs=obs, normalize=True, n_obs=len(obs_t))
    assert Gp["J"] == Gs["J"]
    for key in ("phi", "K1", "K2", "bc"):
        a, b = Gp[key], Gs[key]
        ok = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), ok), key
        scale = np.max(np.abs(b[ok])) if ok.any() else 0.0
        if scale == 0.0:
            assert not np.any(a[ok]), key
            continue
        assert np.max(np.abs(a[ok] - b[ok])) <= 1e-12 * scale, (key, np.max(np.abs(a[ok] - b[ok])) / scale)
