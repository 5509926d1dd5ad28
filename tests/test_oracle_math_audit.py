"""Where the kernels' arithmetic can part from the reference's: only in a
transcendental result.  The oracle is built twice from the same restatement
of the reference's loops, with glibc's exp / log / cbrt (the reference's) and
with the kernels' (csrc/sw_math.h: glibc's exp and cbrt, a correctly rounded
log), the second auditing every call against glibc (oracle/orc_math.h).

On every golden case -- native maps, the _E dry runs of the 4D-Var layer
(the MUSCL-EVR ones whose limiter amplifies any ulp) and the reference's
functional case sp_toce -- the two builds agree bit for bit except where the
audit found a differing result first: a state can only differ after a step
in which a transcendental call returned a different double.  In all but one
case they agree to the last bit over the full run, although glibc's log
misrounds some arguments (the following exp absorbs those); the exception,
rain_mixed_godunov_infil_coeff_everywhere, moves by 1.5e-15 from its first
differing log on."""
import os

import numpy as np
import pytest

from conftest import (e_golden_names, golden_names, load_golden, load_sp_toce, map_time, regen_case,
                      rel_err)


def test_kernel_math_equals_glibc_exp_cbrt_and_rounds_log_correctly(oracle):
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-30, 30, 200_000), rng.uniform(-745, 709, 200_000),
                        np.exp(rng.uniform(-700, 700, 200_000))])
    for fn in ("exp", "cbrt"):
        d, g = oracle.math_eval(fn, x, "device"), oracle.math_eval(fn, x, "glibc")
        assert np.array_equal(d.view(np.uint64), g.view(np.uint64)), fn
    xp = np.concatenate([np.exp(rng.uniform(-700, 700, 300_000)), rng.uniform(0.5, 2.0, 300_000)])
    d, g = oracle.math_eval("log", xp, "device"), oracle.math_eval("log", xp, "glibc")
    diff = d != g
    # glibc's log is not correctly rounded on ~1e-4..1e-3 of these; where the
    # two differ, the kernels' result is the nearer (long double reference)
    assert 0 < diff.mean() < 2e-3
    ref = np.log(xp[diff].astype(np.longdouble))
    assert np.all(np.abs(d[diff] - ref) <= np.abs(g[diff] - ref))


@pytest.mark.parametrize("name", [n for n in golden_names() if not n.endswith("_error")])
def test_native_builds_part_only_after_a_differing_transcendental(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    A = oracle.OracleSolver(case, math="glibc")
    oracle.math_audit_reset()
    B = oracle.OracleSolver(case, math="device")
    t_end = map_time(g, len(g["files"]) - 1)
    first_diff = None
    n = 0
    while A.t < t_end:
        A.step(1)
        B.step(1)
        if first_diff is None and any(not np.array_equal(a, b) for a, b in zip(A.state(), B.state())):
            first_diff = n
        n += 1
    au = oracle.math_audit()
    if first_diff is not None:
        assert 0 <= au["step"] <= first_diff, (au, first_diff)
        for a, b in zip(A.state(), B.state()):
            assert rel_err(b, a) <= 1e-12
    expect_exact = name != "rain_mixed_godunov_infil_coeff_everywhere"
    assert (first_diff is None) == expect_exact, (name, first_diff, au)


@pytest.mark.parametrize("name", e_golden_names())
def test_e_dry_runs_are_the_reference_with_the_kernel_math(sw, oracle, tmp_path, name):
    g = regen_case(sw, name, str(tmp_path))
    case = sw.Case(str(tmp_path))
    oracle.math_audit_reset()
    r = oracle.OracleE(case, math="device").record(list(g["obs_t"]), max_steps=20000, states=True)
    au = oracle.math_audit()
    np.testing.assert_array_equal(r["dt"], g["dt"])
    np.testing.assert_array_equal(r["trig"], g["trig"])
    last = r["states"][-1]  # [field][cell] after the last step
    for k, key in enumerate(("h", "qx", "qy")):
        np.testing.assert_array_equal(last[k], g[key])
    # the inflow cases do meet log arguments glibc misrounds: absorbed
    if name.startswith("e_scale"):
        assert au["count"] > 0 and au["fn"] == "log"


def test_sp_toce_dry_run_is_the_reference_with_the_kernel_math(sw, oracle):
    case, g = load_sp_toce(sw)
    ones = np.ones(case.n_cell)
    r0 = oracle.OracleE(case).record(g["obs_t"], phi=ones, max_steps=20000)
    oracle.math_audit_reset()
    r1 = oracle.OracleE(case, math="device").record(g["obs_t"], phi=ones, max_steps=20000)
    au = oracle.math_audit()
    assert len(r0["dt"]) == len(r1["dt"]) == 3157
    np.testing.assert_array_equal(r0["dt"], r1["dt"])
    assert au["count"] > 0  # 36 log calls differ from glibc's; none changes a state
