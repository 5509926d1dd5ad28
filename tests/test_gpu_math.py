"""The device's exp, log and cbrt (csrc/sw_math.h, through sw2dg_math_eval)
are bit-identical to a host build of the same code (the device-math oracle,
oracle/orc_math.h), and through it exp and cbrt to glibc's -- the functions
the reference runs on -- on 3e6 arguments over the ranges the solver meets
(friction and inflow arguments, depths, wide exponents, special values)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _args(rng, n):
    return np.concatenate([
        rng.uniform(-30.0, 30.0, n), rng.uniform(-745.0, 709.0, n),
        np.exp(rng.uniform(-700.0, 700.0, n)), rng.uniform(0.9, 1.1, n),
        rng.uniform(0.0, 1e-3, n), -rng.uniform(0.0, 5.0, n),
        np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 5e-324, 1e-310, 1e308, -745.2,
                  709.8, 1e-20])])


@pytest.mark.parametrize("fn", ["exp", "log", "cbrt"])
def test_gpu_math_is_bitwise_the_host_build(sw, oracle, fn):
    from sw2d_amd import _native as N
    rng = np.random.default_rng({"exp": 1, "log": 2, "cbrt": 3}[fn])
    x = _args(rng, 500_000)
    y = np.empty_like(x)
    print(f"[gpu] sw_math {fn} on {len(x)} values", flush=True)
    rc = N.gpu_lib().sw2dg_math_eval({"exp": 0, "log": 1, "cbrt": 2}[fn], len(x), N.dptr(x), N.dptr(y), 0)
    assert rc == 0
    host = oracle.math_eval(fn, x, math="device")
    assert np.array_equal(y.view(np.uint64)[~np.isnan(host)], host.view(np.uint64)[~np.isnan(host)])
    assert np.array_equal(np.isnan(y), np.isnan(host))
    if fn != "log":  # exp and cbrt are glibc's
        g = oracle.math_eval(fn, x, math="glibc")
        ok = ~np.isnan(g)
        assert np.array_equal(y.view(np.uint64)[ok], g.view(np.uint64)[ok])
