"""Spalart-Allmaras (sixth equation, coupled) on the GPU against the CPU
restatement, and the tangent/adjoint/Newton machinery at NV = 6.

- residual: plate with prisms, SA, VK-limited curvilinear/lsqr gradients,
  noisy state, at the residual bar (1e-12 of the per-cell summation scale);
- GPU wall distances (forced on a small mesh) equal the host brute force
  bit for bit, cells and faces;
- FD Jacobian incl. the SA source block vs the oracle's;
- FULL tangent vs central differences of the GPU residual (smooth state, no
  limiter), and the dot-product test against the adjoint;
- Newton steps (FD/FP64 and JVP/FP32) reduce the residual.
"""
import os

import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu


def _plate(noise=0.01, **kw):
    return cases.plate_case(12, 8, 3, first=1e-3, noise=noise, sa=True, **kw)


def test_sa_residual_parity():
    from test_residual_gpu import _compare
    _compare(_plate(), label="SA plate")


def test_sa_residual_parity_first_order_lsqr():
    from test_residual_gpu import _compare
    cs = _plate()
    cs.opt.order = A.ORDER_FIRST
    cs.opt.grad_hex = cs.opt.grad_prism = A.GRAD_LSQR
    _compare(cs, label="SA plate first order lsqr")


def test_gpu_wall_distance_matches_host(monkeypatch):
    from freecfd_amd.ns import Grid, NavierStokes
    cs = _plate()
    ref = Grid(cs.raw, cs.opt).metrics()["wall_distance"]
    monkeypatch.setenv("FCFD_HOST_WALL_DISTANCE_MAX_CELLS", "0")
    ns = NavierStokes(cs.raw, cs.opt)
    got = ns.metrics()["wall_distance"]
    assert np.array_equal(got, ref)
    # residual parity with the device-computed distances (SA source, SW dissipation)
    cs.opt.convective_flux = A.FLUX_SW
    from test_residual_gpu import _compare
    _compare(cs, label="SA plate SW, device wall distances")


def test_sa_fd_jacobian_parity():
    from test_jacobian_gpu import _compare
    _compare(_plate(), label="SA plate", tol=1e-6)


def _smooth_plate():
    cs = cases.smooth_state(_plate(noise=0.0))
    cs.opt.limiter = A.LIMITER_NONE
    return cs


def test_sa_full_jvp_vs_central_fd():
    from test_tangent_gpu import _fd_check
    rel = _fd_check(_smooth_plate(), "SA, no limiter")
    assert rel.max() < 1e-6


def test_sa_dot_product():
    from test_adjoint_gpu import _dot_check
    assert _dot_check(_plate()) < 1e-11
    cs = _plate()
    cs.opt.limiter = A.LIMITER_BJ
    assert _dot_check(cs) < 1e-11


@pytest.mark.parametrize("jac,fp32", [(A.JAC_FD, False), (A.JAC_JVP_FULL, True)])
def test_sa_newton_steps_reduce_residual(jac, fp32):
    from freecfd_amd.ns import NavierStokes
    cs = _plate()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    hist = []
    for _ in range(8):
        st = ns.newton_step(A.newton_opts(cfl=5.0, jacobian=jac, fp32=fp32, max_iters=30))
        hist.append(st.res_norm)
    print("  ||R||:", " ".join("%.3e" % h for h in hist), " nut norm %.3e" % st.nut_norm, flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.5 * hist[0]
