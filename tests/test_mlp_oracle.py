"""The MLP oracle (oracle/mlp_oracle.py): BF16 rounding known answers and its
backward pass against central finite differences of its own forward."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_rounding_known_answers():
    # BF16 keeps 7 fraction bits: the spacing above 1 is 2^-7
    x = np.array([1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                  1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    y = M.bf16(x)
    # ties go to the even neighbour: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert list(y[:5]) == [1.0, 1.0 + 2.0 ** -7, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7]
    y = y[1:]
    assert y[4] == -2.5
    assert np.isfinite(y[5]) and abs(y[5] - 3.0e38) / 3.0e38 < 2 ** -8
    assert np.isnan(M.bf16(np.array([np.nan], dtype=np.float32)))[0]


def test_oracle_backward_vs_central_fd(sw):
    dims = [12, 16, 16, 5]
    rng = np.random.default_rng(0)
    from sw2d_amd import mlp  # noqa: F401  (init helper lives with the product)
    p = mlp.init_params(dims, 3).astype(np.float64)
    p += 0.05 * rng.standard_normal(p.size)
    x = rng.standard_normal((7, dims[0]))
    t = rng.standard_normal((7, dims[-1]))
    _, g = M.loss_grad(p, dims, x, t, quantise=False)
    for i in rng.choice(p.size, 25, replace=False):
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (M.loss_grad(pp, dims, x, t, False)[0] - M.loss_grad(pm, dims, x, t, False)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
