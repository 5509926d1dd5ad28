"""Host logic of the porosity calibration problem (porosity.py, the mirror
of assim/swSpE_minimize.cpp:267-511) on a stub data term: every penalty's
gradient (ridge, L1, spatial H1, phi_ref) through each change of variable
against central differences of func(), zonation, the reference's grad() J
(no spatial term), and the projected L-BFGS on a bounded quadratic."""
import numpy as np
import pytest


class StubAssim:
    """J_data(m) = sum w (phi(m) - t)^2 with the library's change of variable."""

    def __init__(self, t, w, mode, lam):
        self.t, self.w, self.mode, self.lam = t, w, mode, lam

    def _phi(self, m):
        from sw2d_amd.assim import phi_transform_eval
        lo, hi = (float(m.min()), float(m.max())) if self.mode == "normalize" else (0.0, 0.0)
        return phi_transform_eval(self.mode, m, lam=self.lam, lo=lo, hi=hi)

    def cost(self, phi):
        p, _ = self._phi(np.asarray(phi))
        return float(np.sum(self.w * (p - self.t) ** 2))

    def gradient(self, phi):
        p, d = self._phi(np.asarray(phi))
        return {"J": self.cost(phi), "phi": 2.0 * self.w * (p - self.t) * d}


def _problem(mode, lam=2.0, zonation=False):
    from sw2d_amd.porosity import PorosityProblem
    rng = np.random.default_rng(1)
    n = 40
    phi_b = np.round(rng.uniform(0.3, 0.9, n), 1) if zonation else rng.uniform(0.3, 0.9, n)
    L = np.arange(n - 1)
    R = np.arange(1, n)
    A = StubAssim(rng.uniform(0.2, 0.95, n), rng.uniform(0.5, 1.5, n), mode, lam)
    free = np.sort(rng.choice(n, 30, replace=False))
    P = PorosityProblem(A, phi_b, free_idx=free, alpha=0.7, alpha_l1=0.05, alpha_spatial=0.3,
                        alpha_phi_ref=0.2, phi_ref=rng.uniform(0.3, 0.9, n), zonation=zonation,
                        edges=(L, R), phi_transform=(mode, lam) if mode else None)
    return P, rng


@pytest.mark.parametrize("mode", [None, "tanh", "clamp"])
def test_gradient_matches_differences_of_func(sw, mode):
    P, rng = _problem(mode)
    th = P.theta0() + 0.05 * rng.standard_normal(P.size())
    J, g = P.grad(th)
    m = P.m_of(th)
    p = P.penalties(m)
    assert J == pytest.approx(P.A.cost(m) + p["ridge"] + p["l1"] + p["phi_ref"], rel=1e-14)
    assert P.func(th) == pytest.approx(J + p["spatial"], rel=1e-14)
    e = 1e-7
    for k in range(P.size()):
        d = np.zeros(P.size())
        d[k] = e
        fd = (P.func(th + d) - P.func(th - d)) / (2 * e)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-7), (k, fd, g[k])


def test_zonation_sums_the_cells_of_a_zone(sw):
    P, rng = _problem("tanh", zonation=True)
    assert P.size() == len(np.unique(np.round(P.phi_b[P.free] * 1e6)))
    th = P.theta0() + 0.03 * rng.standard_normal(P.size())
    m = P.m_of(th)
    np.testing.assert_array_equal(m[P.free], th[P.zone_of_free])
    _, g = P.grad(th)
    e = 1e-7
    for k in range(P.size()):
        d = np.zeros(P.size())
        d[k] = e
        fd = (P.func(th + d) - P.func(th - d)) / (2 * e)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-7)


def test_projected_lbfgs_reaches_the_bounded_minimum(sw):
    from sw2d_amd.porosity import minimize
    rng = np.random.default_rng(2)
    n = 12
    c = rng.uniform(-1.0, 2.0, n)
    H = np.diag(rng.uniform(0.5, 3.0, n))
    fg = lambda x: (float(0.5 * (x - c) @ H @ (x - c)), H @ (x - c))
    x, J, tr = minimize(fg, lambda x: fg(x)[0], np.zeros(n), lo=0.0, hi=1.0, eps=1e-10, max_iter=200)
    np.testing.assert_allclose(x, np.clip(c, 0.0, 1.0), atol=1e-8)
