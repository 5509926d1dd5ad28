"""The MLP warm start on the GPU (BF16 MFMA kernels, FP32 accumulation)
against the CPU oracle (oracle/mlp_oracle.py: same BF16 quantisation points,
float64 accumulation), at the network's real shape 2400 -> 512 -> 512 -> 512
-> 68 (20 gauges x 120 samples of 30 s in, 8 Manning zones + 60 hydrograph
knots out).  The two differ by the accumulation precision (FP32 vs float64) and
by the rare BF16 re-quantisation of an activation that lands on the other
side of a rounding boundary, so outputs and gradients are compared
norm-wise at 1e-2 (BF16 resolution 2^-8 = 4e-3 per element)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu
DIMS = [2400, 512, 512, 512, 68]


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def net(sw):
    from sw2d_amd import mlp
    return mlp.MLP(DIMS, max_batch=512, seed=1)


def test_gpu_mlp_forward_matches_oracle(net):
    print("[gpu] mlp forward", flush=True)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((200, DIMS[0])).astype(np.float32)
    y = net.forward(x)
    yo, _, _ = M.forward(net.params(), DIMS, x)
    assert _rel(y, yo) <= 1e-2, _rel(y, yo)
    # a batch that is not a multiple of the 64-row tile, and one larger than the tile grid
    for b in (1, 63, 512):
        xb = rng.standard_normal((b, DIMS[0])).astype(np.float32)
        assert _rel(net.forward(xb), M.forward(net.params(), DIMS, xb)[0]) <= 1e-2


def test_gpu_mlp_gradients_match_oracle(net):
    print("[gpu] mlp gradients", flush=True)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((256, DIMS[0])).astype(np.float32)
    t = rng.standard_normal((256, DIMS[-1])).astype(np.float32)
    loss, g = net.grad(x, t)
    lo, go = M.loss_grad(net.params(), DIMS, x, t)
    assert abs(loss - lo) <= 1e-3 * lo, (loss, lo)
    # layer by layer: W then b
    o = 0
    for a, b in zip(DIMS[:-1], DIMS[1:]):
        for n in (a * b, b):
            assert _rel(g[o: o + n], go[o: o + n]) <= 1e-2, (a, b, n, _rel(g[o: o + n], go[o: o + n]))
            o += n


def test_gpu_mlp_training_fits_a_smooth_map(sw):
    """Adam on the GPU kernels learns a smooth nonlinear map (loss / 20)."""
    from sw2d_amd import mlp
    print("[gpu] mlp training", flush=True)
    rng = np.random.default_rng(2)
    A = rng.standard_normal((DIMS[0], DIMS[-1])) / np.sqrt(DIMS[0])
    x = rng.standard_normal((4096, DIMS[0])).astype(np.float32)
    t = np.tanh(x @ A).astype(np.float32)
    net = mlp.MLP(DIMS, max_batch=256, seed=3)
    l0 = net.grad(x[:256], t[:256])[0]
    hist = net.train(x, t, epochs=20, batch=256, lr=1e-3)
    l1 = net.grad(x[:256], t[:256])[0]
    assert hist[-1] < hist[0] and l1 < l0 / 20, (l0, l1, hist[:3], hist[-3:])


def test_gpu_mlp_resident_epochs_equal_host_fed_steps(sw):
    """sw2dg_mlp_train_epoch (training set resident on the device, one wait
    per epoch) against sw2dg_mlp_train_step per minibatch copied from the
    host: same permutations, parameters and losses bit for bit; a set that
    is not a multiple of the batch (its last partial batch left out)."""
    from sw2d_amd import _native as N
    from sw2d_amd import mlp
    print("[gpu] mlp resident epochs vs host-fed steps", flush=True)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1100, DIMS[0])).astype(np.float32)
    t = rng.standard_normal((1100, DIMS[-1])).astype(np.float32)
    a = mlp.MLP(DIMS, max_batch=256, seed=5)
    b = mlp.MLP(DIMS, max_batch=256, seed=5)
    ha = a.train(x, t, epochs=2, batch=256, lr=1e-3, seed=9, lr_decay=0.9)
    hb = b.train_steps(x, t, epochs=2, batch=256, lr=1e-3, seed=9, lr_decay=0.9)
    assert ha == hb, (ha, hb)
    np.testing.assert_array_equal(a.params(), b.params())
    # one more epoch on the resident set, as warm-start fitting does
    assert a.train_resident(epochs=1, batch=128, lr=5e-4, seed=3) == \
        b.train_steps(x, t, epochs=1, batch=128, lr=5e-4, seed=3)
    np.testing.assert_array_equal(a.params(), b.params())
    rows = np.full(256, 1100, dtype=np.int32)  # outside the set: refused before any launch
    with pytest.raises(N.SW2DError, match="outside"):
        a._check(a.lib.sw2dg_mlp_train_epoch(a.h, 256, 1, rows.ctypes.data_as(N._i32), 1e-3, None))


def test_gpu_mlp_large_batch_tiles_match_small_batches(sw):
    """A batch large enough for the 128 x 128 output tiles (8229 rows: the
    hidden layers' 64 x 64 grid reaches 1032 tiles; the last block of rows
    ragged) gives each row exactly what a small batch gives it: every
    output accumulates its k in the same order whatever the tile."""
    from sw2d_amd import mlp
    print("[gpu] mlp large-batch tiles vs small batches", flush=True)
    net = mlp.MLP(DIMS, max_batch=8229, seed=6)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((8229, DIMS[0])).astype(np.float32)
    y = net.forward(x)
    for lo, hi in ((0, 100), (4000, 4130), (8192, 8229)):
        np.testing.assert_array_equal(y[lo:hi], net.forward(x[lo:hi]))
    assert _rel(y[:64], M.forward(net.params(), DIMS, x[:64])[0]) <= 1e-2
