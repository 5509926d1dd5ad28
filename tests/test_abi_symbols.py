"""The C-ABI libraries load and export every symbol the public headers declare.
No compute call is made (runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

from conftest import PKG_DIR, REPO

HEADERS = [os.path.join(REPO, "include", h) for h in ("sw2dg.h", "sw2d_host.h")]


def declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sw2d[gh]_[A-Za-z0-9_]+)\s*\(", txt)))


def test_headers_declare_something():
    assert len(declared(HEADERS[0])) >= 12
    assert len(declared(HEADERS[1])) >= 15


def test_gpu_library_exports_all_declared_symbols():
    lib = C.CDLL(os.path.join(PKG_DIR, "lib", "libsw2dg.so"))
    missing = [s for h in HEADERS for s in declared(h) if not hasattr(lib, s)]
    assert not missing, missing


def test_host_library_exports_host_symbols():
    lib = C.CDLL(os.path.join(PKG_DIR, "lib", "libsw2d_host.so"))
    missing = [s for s in declared(HEADERS[1]) if not hasattr(lib, s)]
    assert not missing, missing


def test_python_bindings_cover_the_headers(sw):
    N = sw._native
    from sw2d_amd import assim, ensemble, mlp
    assert set(declared(HEADERS[0])) <= (set(N.GPU_SYMBOLS) | set(assim.ASSIM_SYMBOLS)
                                         | set(ensemble.ENS_SYMBOLS) | set(mlp.MLP_SYMBOLS))
    assert set(declared(HEADERS[1])) <= set(N.HOST_SYMBOLS)


def test_version_string_needs_no_device(sw):
    lib = sw._native.gpu_lib()
    assert b"gfx950" in lib.sw2dg_version()


def test_ctx_create_rejects_bad_shapes_before_touching_the_device(sw):
    """Argument validation happens on the host: a negative size is rejected
    with SW2DG_ERR_ARG whether or not a GPU is present."""
    N = sw._native
    lib = N.gpu_lib()
    m = N.Mesh()
    m.n_cell = -1
    p = N.Params()
    out = C.c_void_p()
    assert lib.sw2dg_ctx_create(C.byref(m), C.byref(p), 0, C.byref(out)) == N.ERR_ARG
    assert not out.value
