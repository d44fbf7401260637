"""Adjoint (VJP) kernels and surface loads on the GPU.

- dot-product test <v, J u> == <J^T v, u> against the FULL tangent, for every
  flux, limiter, gradient scheme, wall/inlet/outlet/symmetry BC the cases
  carry, first and second order: both sides are exact derivatives of the
  same branch selections, so they agree to rounding even on noisy states;
- determinism: two VJPs of the same input are bitwise equal;
- surface loads vs the CPU restatement (same face values, same face-order
  sums);
- load adjoint: the VJP with a drag weight on the wall BC, dotted with a
  direction, vs central differences of the GPU loads along it (smooth
  state, no limiter).
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

SCALES = np.array([1e5, 100.0, 100.0, 100.0, 300.0, 1e-4])   # p, u, v, w, T, nu~


def _ns(cs):
    from freecfd_amd.ns import NavierStokes
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    return ns


def _dot_check(cs, seed=11):
    ns = _ns(cs)
    N, nv = cs.n_cells, ns.nv
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(N, nv)) * SCALES[:nv]
    v = rng.normal(size=(N, nv))
    Ju = ns.to_host(ns.jvp(ns.to_device(u), mode=A.JAC_JVP_FULL), nv)
    JTv = ns.to_host(ns.vjp(ns.to_device(v)), nv)
    lhs = float(np.sum(v * Ju))
    rhs = float(np.sum(JTv * u))
    mag = float(np.sum(np.abs(v * Ju)))
    rel = abs(lhs - rhs) / mag
    print("  %s: <v,Ju> = %.15e  <J^T v,u> = %.15e  |diff|/sum|v Ju| = %.2e" % (cs.name, lhs, rhs, rel), flush=True)
    ns.status()
    return rel


def _with(make, **kw):
    def f():
        cs = make()
        for k, val in kw.items():
            setattr(cs.opt, k, val)
        cs.name = cs.name + "".join("_%s%s" % (k, v) for k, v in kw.items())
        return cs
    return f


# Minmod is checked on 3-D meshes only.  On the one-cell-thick wedge the lsqr
# z-gradient is rounding noise, so the minmod ratio (dU + s 1e-12)/(dUg + s 1e-12)
# at the z-symmetry faces has a denominator ~1e-12 and partial derivatives
# ~1e12 w.r.t. the cell, its mirror ghost and its gradient, which cancel in
# exact arithmetic (the ghost mirrors the cell).  The tangent sums them in one
# pass, the adjoint in separate gathers, so the two roundings differ at ~1e-4
# of the product: an ill-conditioned derivative of the reference's minmod on
# 2-D meshes, not a transposition error.
DOT_CASES = [
    ("wedge_bj_roe", _with(lambda: cases.wedge_case(24, 12), limiter=A.LIMITER_BJ)),
    ("wedge_bj_vanleer", _with(lambda: cases.wedge_case(24, 12), limiter=A.LIMITER_BJ, convective_flux=A.FLUX_VAN_LEER)),
    ("wedge_bj_slau", _with(lambda: cases.wedge_case(24, 12), limiter=A.LIMITER_BJ, convective_flux=A.FLUX_SD_SLAU)),
    ("wedge_first_order", _with(lambda: cases.wedge_case(24, 12), order=A.ORDER_FIRST)),
    ("bump_minmod", _with(lambda: cases.bump_case(16, 8, 4, noise=0.01), limiter=A.LIMITER_MINMOD)),
    ("tetbox_minmod_inviscid", _with(lambda: cases.tetbox_case(4, viscous=False), limiter=A.LIMITER_MINMOD)),
    ("cylinder_ausm_sutherland", lambda: cases.cylinder_case(32, 16, noise=0.01)),
    ("tetbox_vk_viscous", lambda: cases.tetbox_case(5)),
    ("bump_curvilinear_vk", lambda: cases.bump_case(16, 8, 4, noise=0.01)),
    ("plate_prisms", lambda: cases.plate_case(12, 8, 3, first=1e-3, noise=0.01)),
    ("plate_sw", _with(lambda: cases.plate_case(12, 8, 3, first=1e-3, noise=0.01), convective_flux=A.FLUX_SW)),
]


@pytest.mark.parametrize("name,make", DOT_CASES)
def test_dot_product(name, make):
    rel = _dot_check(make())
    assert rel < 1e-11


def test_vjp_deterministic():
    cs = cases.tetbox_case(5)
    ns = _ns(cs)
    v = ns.to_device(np.random.default_rng(3).normal(size=(cs.n_cells, 5)))
    a = ns.vjp(v).clone()
    b = ns.vjp(v).clone()
    assert bool((a == b).all())


@pytest.mark.parametrize("name,make", [
    ("wedge", lambda: cases.wedge_case(24, 12)),
    ("cylinder", lambda: cases.cylinder_case(32, 16, noise=0.01)),
    ("plate", lambda: cases.plate_case(12, 8, 3, first=1e-3, noise=0.01)),
    ("bump", lambda: cases.bump_case(16, 8, 4, noise=0.01)),
])
def test_loads_match_oracle(name, make):
    from oracle.oracle import Oracle
    cs = make()
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0)
    ref = o.loads()
    ns = _ns(cs)
    got = ns.loads()
    scale = np.abs(ref).max() + 1e-300
    err = np.abs(got - ref).max() / scale
    print("  %s loads: max |gpu - oracle| / max|load| = %.2e" % (name, err), flush=True)
    print("  ", ref[np.abs(ref).sum(axis=1) > 0], flush=True)
    assert np.abs(ref).max() > 0
    assert err <= 1e-13


@pytest.mark.parametrize("name,make,wall", [
    ("cylinder_viscous", lambda: cases.cylinder_case(32, 16), 0),
    ("plate_viscous", lambda: cases.plate_case(12, 8, 3, first=1e-3), 2),
    ("bump_inviscid", lambda: cases.bump_case(16, 8, 4), 2),
])
def test_load_adjoint_vs_fd(name, make, wall):
    """d(drag)/dW . u from the VJP (drag weight on the wall BC, no residual
    seed) vs central differences of the loads along u."""
    cs = cases.smooth_state(make())
    cs.opt.limiter = A.LIMITER_NONE
    ns = _ns(cs)
    N = cs.n_cells
    e = np.array([1.0, 0.0, 0.0])
    fxw = np.zeros((cs.opt.n_bc, 5))
    fxw[wall, 1:4] = -e        # force = -(D - C)[1..3]
    zero = ns.to_device(np.zeros((N, 5)))
    g = ns.to_host(ns.vjp(zero, fx_weight=fxw), 5)
    u = np.random.default_rng(5).normal(size=(N, 5)) * 1e-7 * np.abs(cs.W0).mean(axis=0)   # ~1e-7 of the state
    adj = float(np.sum(g * u))
    h = 1.0
    ns.set_state(cs.W0 + h * u)
    Fp = ns.loads()[wall, :3] @ e
    ns.set_state(cs.W0 - h * u)
    Fm = ns.loads()[wall, :3] @ e
    fd = (Fp - Fm) / (2 * h)
    rel = abs(adj - fd) / abs(fd)
    print("  %s: adjoint %.10e  central FD %.10e  rel %.2e" % (name, adj, fd, rel), flush=True)
    assert rel < 1e-5
