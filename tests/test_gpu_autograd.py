"""The torch.autograd.Function over the C-ABI gradient: autograd through a
torch parametrisation (Manning n per zone -> K per cell, hydrograph knots ->
boundary table) gives the same 68 gradients as the explicit mapping of the
per-cell / per-step adjoint (whose finite-difference check is in
test_gpu_calibration.py), also when composed with further torch terms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gpu_autograd_matches_adjoint_mapping(sw):
    import torch
    from sw2d_amd import assim, calibration
    from sw2d_amd import _native as N
    from sw2d_amd.autograd import gauge_misfit
    cal = calibration.Calibration(px=40, window=300.0)
    s = sw.SpSolver(cal.case)
    obs = cal.observe(s, cal.n_true, cal.Q_true)
    print("[gpu] autograd", cal.case.n_cell, "cells", flush=True)
    A = assim.Assim(s, cal.obs_t, gauge_cells=cal.gauge_cells, gauge_h=obs)
    # explicit mapping
    cal.apply_hydrograph(s, cal.Q_default)
    K0 = cal.K(cal.n_default)
    G = A.gradient(K1=K0, K2=K0)
    gn, gQ = cal.grad_controls(G, cal.n_default, A.get_record()["t"][:-1])
    # autograd
    n = torch.tensor(cal.n_default, dtype=torch.float64, requires_grad=True)
    Q = torch.tensor(cal.Q_default, dtype=torch.float64, requires_grad=True)
    zone = torch.as_tensor(cal.zone)
    K = 1.0 / n[zone]
    bc_v = torch.stack([torch.zeros_like(Q), Q / cal.west_len])
    types = np.array([N.BC_Q, N.BC_Q], dtype=np.int32)
    J = gauge_misfit(A, K1=K, K2=K, bc_v=bc_v, bc_t=cal.t_knots, bc_types=types)
    J.backward()
    assert float(J.detach()) == G["J"]
    np.testing.assert_allclose(n.grad.numpy(), gn, rtol=1e-9, atol=0)
    np.testing.assert_allclose(Q.grad.numpy(), gQ, rtol=1e-9, atol=1e-300)
    # composed with more torch: the incoming gradient scales the adjoint's
    n.grad, Q.grad = None, None
    K = 1.0 / n[zone]
    J2 = 2.0 * gauge_misfit(A, K1=K, K2=K, bc_v=torch.stack([torch.zeros_like(Q), Q / cal.west_len]),
                            bc_t=cal.t_knots, bc_types=types) + (n ** 2).sum()
    J2.backward()
    np.testing.assert_allclose(n.grad.numpy(), 2.0 * gn + 2.0 * cal.n_default, rtol=1e-9)
    np.testing.assert_allclose(Q.grad.numpy(), 2.0 * gQ, rtol=1e-9, atol=1e-300)
