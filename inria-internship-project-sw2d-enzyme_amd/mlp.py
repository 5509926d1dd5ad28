"""The network of the learned warm start: a dense MLP on the GPU
(include/sw2dg.h, sw2dg_mlp_*), trained and evaluated with the project's
BF16 MFMA kernels (FP32 accumulation, FP32 master weights).

``MLP(dims)`` holds the network; ``forward`` evaluates it, ``grad`` returns
the mean-squared-error loss and its gradient, ``train`` runs Adam over
shuffled minibatches with the training set resident on the device (one
wait per epoch).
"""
import ctypes as C

import numpy as np

from . import _native as N

_f = C.POINTER(C.c_float)
_DECL = {
    "sw2dg_mlp_create": ([C.c_int32, N._i32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
    "sw2dg_mlp_n_params": ([C.c_void_p], C.c_int64),
    "sw2dg_mlp_set_params": ([C.c_void_p, _f], C.c_int),
    "sw2dg_mlp_get_params": ([C.c_void_p, _f], C.c_int),
    "sw2dg_mlp_forward": ([C.c_void_p, C.c_int32, _f, _f], C.c_int),
    "sw2dg_mlp_forward_noalloc": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "sw2dg_mlp_grad": ([C.c_void_p, C.c_int32, _f, _f, _f, _f], C.c_int),
    "sw2dg_mlp_train_step": ([C.c_void_p, C.c_int32, _f, _f, C.c_float, _f], C.c_int),
    "sw2dg_mlp_set_data": ([C.c_void_p, C.c_int64, _f, _f], C.c_int),
    "sw2dg_mlp_train_epoch": ([C.c_void_p, C.c_int32, C.c_int64, N._i32, C.c_float, _f], C.c_int),
    "sw2dg_mlp_last_error": ([C.c_void_p], C.c_char_p),
    "sw2dg_mlp_destroy": ([C.c_void_p], None),
}
MLP_SYMBOLS = tuple(_DECL)


def _lib():
    lib = N.gpu_lib()
    if not getattr(lib, "_mlp_ready", False):
        for k, (a, r) in _DECL.items():
            f = getattr(lib, k)
            f.argtypes = a
            f.restype = r
        lib._mlp_ready = True
    return lib


def _fp(a):
    return a.ctypes.data_as(_f)


def init_params(dims, seed):
    """He-style init, flattened layer by layer (W [in x out] then b)."""
    rng = np.random.default_rng(seed)
    out = []
    for a, b in zip(dims[:-1], dims[1:]):
        out.append((rng.standard_normal((a, b)) * np.sqrt(2.0 / a)).astype(np.float32).ravel())
        out.append(np.zeros(b, dtype=np.float32))
    return np.concatenate(out)


class MLP:
    def __init__(self, dims, max_batch=1024, seed=0, device=-1):
        self.lib = _lib()
        self.dims = [int(d) for d in dims]
        self.max_batch = int(max_batch)
        d = np.ascontiguousarray(self.dims, dtype=np.int32)
        h = C.c_void_p()
        rc = self.lib.sw2dg_mlp_create(len(self.dims) - 1, N.i32ptr(d), self.max_batch, device, C.byref(h))
        if rc != 0:
            raise N.SW2DError(rc, "sw2dg_mlp_create failed")
        self.h = h
        self.n_params = int(self.lib.sw2dg_mlp_n_params(h))
        self.set_params(init_params(self.dims, seed))

    def _check(self, rc):
        if rc != 0:
            raise N.SW2DError(rc, self.lib.sw2dg_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.sw2dg_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._check(self.lib.sw2dg_mlp_set_params(self.h, _fp(p)))

    def params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.sw2dg_mlp_get_params(self.h, _fp(p)))
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = []
        for i in range(0, len(x), self.max_batch):
            xb = np.ascontiguousarray(x[i: i + self.max_batch])
            y = np.empty((len(xb), self.dims[-1]), dtype=np.float32)
            self._check(self.lib.sw2dg_mlp_forward(self.h, len(xb), _fp(xb), _fp(y)))
            out.append(y)
        return np.concatenate(out)

    def forward_noalloc(self, x_dev, y_dev, batch, stream=None):
        """sw2dg_mlp_forward_noalloc: y = MLP(x) for `batch` rows with x and y
        given as device addresses (ints, e.g. a tensor's data_ptr()); nothing
        is allocated.  stream: a hipStream_t handle to enqueue on without
        waiting, or None for the network's own stream (waited for)."""
        self._check(self.lib.sw2dg_mlp_forward_noalloc(self.h, int(batch), C.c_void_p(int(x_dev)),
                                                       C.c_void_p(int(y_dev)),
                                                       None if stream is None else C.c_void_p(int(stream))))

    def grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.sw2dg_mlp_grad(self.h, len(x), _fp(x), _fp(t), C.byref(loss), _fp(g)))
        return float(loss.value), g

    def set_data(self, x, t):
        """Keeps the training set (x [n x dims[0]], t [n x dims[-1]]) on the device."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        if x.ndim != 2 or t.ndim != 2 or len(x) != len(t) or x.shape[1] != self.dims[0] \
                or t.shape[1] != self.dims[-1]:
            raise ValueError(f"training set shapes {x.shape}, {t.shape} do not fit dims {self.dims}")
        self._check(self.lib.sw2dg_mlp_set_data(self.h, len(x), _fp(x), _fp(t)))
        self.n_data = len(x)

    def train(self, x, t, epochs=50, batch=256, lr=1e-3, seed=0, lr_decay=1.0):
        """Adam over shuffled minibatches (each epoch a fresh permutation,
        the last partial batch left out); returns the mean loss of every
        epoch.  The set goes to the device once (set_data), then
        train_resident."""
        self.set_data(x, t)
        return self.train_resident(epochs=epochs, batch=batch, lr=lr, seed=seed, lr_decay=lr_decay)

    def train_resident(self, epochs=50, batch=256, lr=1e-3, seed=0, lr_decay=1.0):
        """``train`` on the set last given to set_data: one
        sw2dg_mlp_train_epoch call (one wait) per epoch."""
        n = getattr(self, "n_data", 0)
        if n == 0:
            raise ValueError("no training set on the device: call set_data first")
        rng = np.random.default_rng(seed)
        n_steps = n // batch
        hist = []
        for ep in range(epochs):
            idx = rng.permutation(n)
            tot = 0.0
            if n_steps:
                rows = np.ascontiguousarray(idx[: n_steps * batch], dtype=np.int32)
                losses = np.empty(n_steps, dtype=np.float32)
                self._check(self.lib.sw2dg_mlp_train_epoch(self.h, int(batch), n_steps,
                                                           rows.ctypes.data_as(N._i32), float(lr),
                                                           _fp(losses)))
                for v in losses:
                    tot += float(v)
            hist.append(tot / max(n_steps, 1))
            lr *= lr_decay
        return hist

    def train_steps(self, x, t, epochs=1, batch=256, lr=1e-3, seed=0, lr_decay=1.0):
        """``train`` one sw2dg_mlp_train_step call per minibatch (the batch
        copied from the host each step); the same results."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        rng = np.random.default_rng(seed)
        hist = []
        loss = C.c_float()
        for ep in range(epochs):
            idx = rng.permutation(len(x))
            tot, nb = 0.0, 0
            for i in range(0, len(x) - batch + 1, batch):
                sel = idx[i: i + batch]
                xb = np.ascontiguousarray(x[sel])
                tb = np.ascontiguousarray(t[sel])
                self._check(self.lib.sw2dg_mlp_train_step(self.h, len(sel), _fp(xb), _fp(tb),
                                                          float(lr), C.byref(loss)))
                tot += loss.value
                nb += 1
            hist.append(tot / max(nb, 1))
            lr *= lr_decay
        return hist

