"""sw2dg_mlp_forward_noalloc (include/sw2dg.h): the learned warm start's
forward pass with its activations in the library's fixed device array, for
callers that cannot allocate.  Checked through the C-ABI against
sw2dg_mlp_forward (same kernels, so bit for bit), on consecutive calls whose
second is smaller (the array keeps the first call's contents), on a batch
above max_batch, on a caller's stream, and the refusal of a batch the array
cannot hold.

The test's own device buffers come from the HIP runtime the library is
linked against (libamdhip64.so.7, already loaded with it), not from torch,
which bundles a second HIP runtime that cannot open the GPU once the first
one has."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [2400, 512, 512, 512, 96]  # the warm start's widths: 20 x 120 gauge inputs
H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]),
                           ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipStreamCreate", [C.POINTER(C.c_void_p)]),
                           ("hipStreamSynchronize", [C.c_void_p]),
                           ("hipStreamDestroy", [C.c_void_p]),
                           ("hipDeviceSynchronize", [])):
            f = getattr(self.rt, name)
            f.argtypes, f.restype = args, C.c_int
        self.held = []

    def ok(self, rc, what):
        assert rc == 0, f"{what}: HIP error {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        self.held.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self.ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def get(self, p, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        self.ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def close(self):
        self.rt.hipDeviceSynchronize()
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


@pytest.fixture
def hip(sw):
    from sw2d_amd import _native as N
    N.gpu_lib()  # the library first: its HIP runtime is the one used below
    h = Hip()
    yield h
    h.close()


def test_gpu_mlp_forward_noalloc_equals_forward(hip):
    from sw2d_amd.mlp import MLP
    print("[gpu] mlp forward_noalloc vs forward", flush=True)
    net = MLP(DIMS, max_batch=512, seed=7)
    rng = np.random.default_rng(11)
    for B in (384, 29, 1500):  # the second call smaller than the first; the third above max_batch
        x = rng.standard_normal((B, DIMS[0])).astype(np.float32)
        y_ref = net.forward(x)  # sw2dg_mlp_forward, in chunks of max_batch rows
        dx = hip.put(x)
        dy = hip.put(np.full((B, DIMS[-1]), np.nan, dtype=np.float32))
        net.forward_noalloc(dx, dy, B)
        y = hip.get(dy, (B, DIMS[-1]))
        assert np.isfinite(y).all()
        np.testing.assert_array_equal(y, y_ref)
    net.close()


def test_gpu_mlp_forward_noalloc_on_a_caller_stream_and_its_limit(hip):
    from sw2d_amd import _native as N
    from sw2d_amd.mlp import MLP
    print("[gpu] mlp forward_noalloc on a caller's stream; size limit", flush=True)
    net = MLP(DIMS, max_batch=256, seed=3)
    rng = np.random.default_rng(5)
    B = 200
    x = rng.standard_normal((B, DIMS[0])).astype(np.float32)
    y_ref = net.forward(x)
    dx = hip.put(x)
    dy = hip.put(np.zeros((B, DIMS[-1]), dtype=np.float32))
    st = C.c_void_p()
    hip.ok(hip.rt.hipStreamCreate(C.byref(st)), "hipStreamCreate")
    try:
        net.forward_noalloc(dx, dy, B, stream=st.value)
        hip.ok(hip.rt.hipStreamSynchronize(st), "hipStreamSynchronize")
    finally:
        hip.rt.hipStreamDestroy(st)
    np.testing.assert_array_equal(hip.get(dy, (B, DIMS[-1])), y_ref)
    # a batch whose activations exceed the 64 MiB array is refused before any launch
    per_row = 2 * (DIMS[0] + sum(DIMS[1:-1]))
    too_big = (64 << 20) // per_row + 1
    with pytest.raises(N.SW2DError, match="maximum"):
        net.forward_noalloc(dx, dy, too_big)
    net.close()
