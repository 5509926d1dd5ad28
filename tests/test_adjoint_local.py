"""Hand-written reverse-mode derivatives of the local _E functions against
central finite differences (dot-product form: <J^T w, v> = d/de <w, f(x+e v)>).

Inputs are drawn in smooth regions (away from the regime switches) so the
finite differences see one branch.  A planted error (a derivative with a
dropped term) must fail the same check, so the check has teeth."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG_DIR

LIB = C.CDLL(os.path.join(PKG_DIR, "lib", "libsw2d_adjtest.so"))
dp = C.POINTER(C.c_double)
HMIN, PHIMIN = 1e-5, 1e-16


def arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def fwd(name, p, nout, *extra):
    out = np.zeros(nout)
    getattr(LIB, name)(arr(p).ctypes.data_as(dp), *extra, out.ctypes.data_as(dp))
    return out


def vjp(name, p, w, nin, *extra):
    d = np.zeros(nin)
    getattr(LIB, name)(arr(p).ctypes.data_as(dp), *extra, arr(w).ctypes.data_as(dp),
                       d.ctypes.data_as(dp))
    return d


def check(f, g, x, active, nout, rng, eps=1e-6, tol=1e-6, trials=5):
    """f(x) -> outputs; g(x, w) -> adjoints of the active inputs."""
    worst = 0.0
    for _ in range(trials):
        w = rng.standard_normal(nout)
        v = np.zeros_like(x)
        v[active] = rng.standard_normal(len(active))
        scale = np.maximum(np.abs(x), 1.0)
        vx = v * scale
        fd = (w @ f(x + eps * vx) - w @ f(x - eps * vx)) / (2 * eps)
        ad = g(x, w) @ vx[active]
        err = abs(fd - ad) / max(abs(fd), abs(ad), 1e-12)
        worst = max(worst, err)
    return worst


def test_lr_state(rng=np.random.default_rng(0)):
    for _ in range(20):
        h = rng.uniform(0.05, 3.0)
        x = np.array([h, rng.normal(0, 1), rng.normal(0, 1), rng.uniform(-1, 1), 1.0, 0.6, 0.8])
        e = check(lambda p: fwd("sw2dt_lr_fwd", p, 5, C.c_double(HMIN)),
                  lambda p, w: vjp("sw2dt_lr_vjp", p, w, 3, C.c_double(HMIN)),
                  x, [0, 1, 2], 5, rng)
        assert e < 1e-6, e


def _hllc_point(rng, regime="both"):
    bL = bR = 1.0
    zbL, zbR = rng.uniform(0, 0.5, 2)
    hL, hR = rng.uniform(0.3, 2.0, 2)
    if regime == "rhs":
        hL = 0.0
        zbL = zbR + hR + 0.5
    phiL, phiR = rng.uniform(0.4, 1.0, 2)
    qL, rL, qR, rR = rng.normal(0, 0.5, 4)
    lam = lambda h, q: (q / h - np.sqrt(9.81 * h), q / h + np.sqrt(9.81 * h)) if h > 0 else (0.0, 0.0)
    lmL, lpL = lam(hL, qL)
    lmR, lpR = lam(hR, qR)
    return np.array([bL, bR, phiL, phiR, zbL, zbR, zbL + hL, qL, rL, zbR + hR, qR, rR, lmL, lmR, lpL, lpR])


@pytest.mark.parametrize("regime", ["both", "rhs"])
def test_hllc(regime):
    rng = np.random.default_rng(1)
    active = [6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 2, 3]  # matches the adjoint's d[0..11] order
    for _ in range(30):
        x = _hllc_point(rng, regime)
        f = lambda p: fwd("sw2dt_hllc_fwd", p, 5, C.c_double(HMIN), C.c_double(PHIMIN))
        g = lambda p, w: vjp("sw2dt_hllc_vjp", p, w, 12, C.c_double(HMIN), C.c_double(PHIMIN))
        assert check(f, g, x, active, 5, rng) < 2e-6


def test_hllc_planted_error_is_detected():
    """Drop the porosity derivative of the smoothed min: the check must fail."""
    rng = np.random.default_rng(2)
    x = _hllc_point(rng)
    active = [6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 2, 3]
    f = lambda p: fwd("sw2dt_hllc_fwd", p, 5, C.c_double(HMIN), C.c_double(PHIMIN))

    def g_bad(p, w):
        d = vjp("sw2dt_hllc_vjp", p, w, 12, C.c_double(HMIN), C.c_double(PHIMIN))
        d[10] = 0.0
        return d

    assert check(f, g_bad, x, active, 5, rng) > 1e-3


@pytest.mark.parametrize("btype,bc,h", [(3, 0.8, 0.02), (3, 0.8, 1.5), (0, 0.65, 1.2), (1, 1.7, 1.2),
                                        (2, 2.4, 1.0), (1, 1.1, 0.0)])
def test_boundary_flux(btype, bc, h):
    rng = np.random.default_rng(3)
    for _ in range(10):
        zb = rng.uniform(0, 0.3)
        q, r = rng.normal(0, 0.3, 2)
        u = q / h if h > HMIN else 0.0
        lp = u + np.sqrt(9.81 * h)
        x = np.array([1.0, rng.uniform(0.5, 1.0), zb, zb + h, q, r, lp, bc])
        f = lambda p: fwd("sw2dt_bnd_fwd", p, 3, C.c_int(btype), C.c_double(HMIN), C.c_double(PHIMIN))
        g = lambda p, w: vjp("sw2dt_bnd_vjp", p, w, 6, C.c_int(btype), C.c_double(HMIN), C.c_double(PHIMIN))
        active = [3, 4, 5, 6, 1, 7] if h > 0 else [4, 5, 6, 1, 7]
        if h == 0:
            g2 = lambda p, w: g(p, w)[[1, 2, 3, 4, 5]]
            assert check(f, g2, x, active, 3, rng, eps=1e-7) < 2e-6
        else:
            assert check(f, g, x, active, 3, rng) < 2e-6


def test_friction():
    rng = np.random.default_rng(4)
    for _ in range(20):
        a = rng.uniform(-0.5, 0.5)
        x = np.array([rng.uniform(0.1, 3), rng.normal(0, 1), rng.normal(0, 1), np.cos(a), np.sin(a),
                      rng.uniform(15, 60), rng.uniform(15, 60), rng.uniform(0.01, 1.0)])
        f = lambda p: fwd("sw2dt_fric_fwd", p, 2, C.c_double(HMIN))
        g = lambda p, w: vjp("sw2dt_fric_vjp", p, w, 5, C.c_double(HMIN))
        assert check(f, g, x, [0, 1, 2, 5, 6], 2, rng) < 2e-6


def test_head_loss():
    rng = np.random.default_rng(5)
    for _ in range(20):
        a = rng.uniform(-0.5, 0.5)
        x = np.array([rng.uniform(0.2, 3), rng.normal(0, 1), rng.normal(0, 1), np.cos(a), np.sin(a),
                      0.05, rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0.01, 0.5)])
        f = lambda p: fwd("sw2dt_hl_fwd", p, 2, C.c_double(HMIN))
        g = lambda p, w: vjp("sw2dt_hl_vjp", p, w, 3, C.c_double(HMIN))
        assert check(f, g, x, [0, 1, 2], 2, rng) < 2e-6


def _recon_point(rng, side, h=None):
    h = rng.uniform(0.05, 2.0) if h is None else h
    zb = rng.uniform(0.0, 0.5)
    nx, ny = np.cos(a := rng.uniform(0, 2 * np.pi)), np.sin(a)
    sx, sy = rng.normal(0, 0.05, 2)
    dxm, dym = rng.normal(0, 2.0, 2)
    # the midpoint lies on the side of the normal that the cell's side implies
    return np.array([h, rng.normal(0, 0.5), rng.normal(0, 0.5), zb, 1.0, sx, sy, nx, ny, dxm, dym,
                     rng.uniform(0.05, 0.5)])


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("h", [None, 3e-3])  # wet, and below the desingularisation depth
def test_muscl_reconstruction(side, h):
    """recon_muscl (one side of the MUSCL-EVR internal-edge states) against
    its VJP: h, qx, qy and the limited slope (sx, sy) are active."""
    rng = np.random.default_rng(10 + side)
    f = lambda p: fwd("sw2dt_recon_fwd", p, 5, C.c_int(side))
    g = lambda p, w: vjp("sw2dt_recon_vjp", p, w, 5, C.c_int(side))
    tested = 0
    for _ in range(40):
        x = _recon_point(rng, side, h)
        if h is not None:
            x[1:3] *= 1e-3
            x[5:7] *= 1e-2
        # keep away from a reconstructed depth near 0, where sqrt(g h) bends
        # too sharply for central differences
        if f(x)[0] - x[3] < 0.2 * x[0]:
            continue
        tested += 1
        assert check(f, g, x, [0, 1, 2, 5, 6], 5, rng, eps=1e-7) < 2e-6
    assert tested >= 25


def test_muscl_boundary_state():
    rng = np.random.default_rng(12)
    f = lambda p: fwd("sw2dt_bnd_muscl_fwd", p, 4)
    g = lambda p, w: vjp("sw2dt_bnd_muscl_vjp", p, w, 3)
    for h in (1e-3, 0.02, 1.5):
        for _ in range(10):
            a = rng.uniform(0, 2 * np.pi)
            x = np.array([h, rng.normal(0, 0.5) * h, rng.normal(0, 0.5) * h, 0.1, 1.0, np.cos(a), np.sin(a)])
            assert check(f, g, x, [0, 1, 2], 4, rng, eps=1e-7) < 2e-6


def _slope_cell(rng, deg):
    """A cell at the origin with deg neighbours around it, midpoints halfway,
    free surfaces varied enough that the limiter scales most slopes."""
    ang = np.sort(rng.uniform(0, 2 * np.pi, deg))
    r = rng.uniform(0.5, 1.5, deg)
    xj, yj = r * np.cos(ang), r * np.sin(ang)
    nb = np.zeros((deg, 6))
    nb[:, 1], nb[:, 2] = xj, yj
    nb[:, 3], nb[:, 4] = 0.55 * xj, 0.55 * yj
    nb[:, 5] = rng.integers(0, 2, deg)
    z = rng.normal(0, 1.0, deg + 1)
    return nb, z


def _slope_f(nb, deg):
    def f(p):
        n2 = nb.copy()
        n2[:, 0] = p[1:]
        out = np.zeros(2)
        LIB.sw2dt_slope_fwd(arr([p[0], 0.0, 0.0]).ctypes.data_as(dp), arr(n2).ctypes.data_as(dp),
                            C.c_int(deg), out.ctypes.data_as(dp))
        return out
    return f


def _slope_g(nb, deg):
    def g(p, w):
        n2 = nb.copy()
        n2[:, 0] = p[1:]
        d = np.zeros(deg + 1)
        LIB.sw2dt_slope_vjp(arr([p[0], 0.0, 0.0]).ctypes.data_as(dp), arr(n2).ctypes.data_as(dp),
                            C.c_int(deg), arr(w).ctypes.data_as(dp), d.ctypes.data_as(dp))
        return d
    return g


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 6])
def test_muscl_slope_and_limiter(deg):
    """The limited free-surface slope of one cell against its VJP, with the
    free surfaces of the cell and of every neighbour active."""
    rng = np.random.default_rng(20 + deg)
    scaled = 0
    for _ in range(40):
        nb, z = _slope_cell(rng, deg)
        f, g = _slope_f(nb, deg), _slope_g(nb, deg)
        e = check(f, g, z, list(range(deg + 1)), 2, rng, eps=1e-7, trials=3)
        assert e < 2e-6 or deg == 1, e
        if deg > 1:  # count cells whose limiter acted
            scaled += not np.allclose(f(z), _raw_slope(nb, z))
    assert deg <= 1 or scaled >= 10, scaled


def _raw_slope(nb, z):
    dx, dy = nb[:, 1], nb[:, 2]
    d2 = dx * dx + dy * dy
    w = 1.0 / np.sqrt(d2)
    return np.array([np.sum((z[1:] - z[0]) * dx / d2 * w), np.sum((z[1:] - z[0]) * dy / d2 * w)]) / np.sum(w)


def test_muscl_slope_planted_error_is_detected():
    """Hold the limiter's scale factor constant in the reverse (its
    derivative dropped): the check must fail on cells where it acts."""
    rng = np.random.default_rng(30)
    worst = 0.0
    for _ in range(20):
        nb, z = _slope_cell(rng, 4)
        f = _slope_f(nb, 4)

        def g_bad(p, w, nb=nb):
            # the scaled slope reversed as if every visit's sf were frozen:
            # sf^T w on the raw slope's linear map, by finite differences of
            # the raw slope alone
            s_lim, s_raw = f(p), _raw_slope(nb, p)
            ratio = np.divide(s_lim, s_raw, out=np.ones(2), where=np.abs(s_raw) > 1e-300)
            J = np.zeros((2, 5))
            for k in range(5):
                v = np.zeros(5)
                v[k] = 1e-7
                J[:, k] = (_raw_slope(nb, p + v) - _raw_slope(nb, p - v)) / 2e-7
            return (w * ratio) @ J

        worst = max(worst, check(f, g_bad, z, list(range(5)), 2, rng, eps=1e-7, trials=2))
    assert worst > 1e-2, worst


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 6])
def test_muscl_slope_register_variant_equals_replay(deg):
    """slope_vjp_reg (incidences and visit slopes kept in registers, the
    kernels' path up to 4 internal edges) against slope_vjp (visits
    replayed): the same adjoints to round-off, and none beyond its bound."""
    rng = np.random.default_rng(40 + deg)
    for _ in range(30):
        nb, z = _slope_cell(rng, deg)
        w = rng.standard_normal(2)
        d1 = _slope_g(nb, deg)(z, w)
        n2 = nb.copy()
        n2[:, 0] = z[1:]
        d2 = np.zeros(deg + 1)
        ok = LIB.sw2dt_slope_vjp_reg(arr([z[0], 0.0, 0.0]).ctypes.data_as(dp), arr(n2).ctypes.data_as(dp),
                                     C.c_int(deg), arr(w).ctypes.data_as(dp), d2.ctypes.data_as(dp))
        assert ok == 1
        np.testing.assert_allclose(d2, d1, rtol=1e-13, atol=1e-13 * np.max(np.abs(d1)))
    nb, z = _slope_cell(rng, 7)
    d = np.zeros(8)
    assert LIB.sw2dt_slope_vjp_reg(arr([z[0], 0.0, 0.0]).ctypes.data_as(dp), arr(nb).ctypes.data_as(dp),
                                   C.c_int(7), arr([1.0, 1.0]).ctypes.data_as(dp), d.ctypes.data_as(dp)) == 0
