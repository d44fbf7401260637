"""Tangent (JVP) kernels on the GPU.

- FULL tangent vs central finite differences of the GPU residual itself
  (second-order residual with gradients, limiters and ghosts); the direction
  is scaled per variable so that h*v is ~1e-7 of the state, as in the
  differentiation-check configuration;
- FROZEN tangent vs the assembled reference FD Jacobian of the CPU
  restatement applied to the same direction (agreement to the FD truncation);
- exact linearity: J(2u) == 2 J(u) bitwise (every derivative operation is
  linear in the tangents and doubling is exact), J(u+v) = Ju + Jv to rounding;
- first-order inviscid: FULL == FROZEN (no gradients or ghost centres enter).
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


def _direction(n, seed, nv=5):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, nv)) * SCALES[:nv]


def _residual_at(ns, W_host):
    ns.set_state(W_host)
    ns.assemble_residual()
    return ns.residual_host()


def _smooth(cs, seed=21, amp=0.02):
    return cases.smooth_state(cs, seed, amp)


def _fd_check(cs, label):
    print("case", cs.name, label, flush=True)
    ns = _ns(cs)
    v = _direction(cs.n_cells, 4, ns.nv)
    Jv = ns.to_host(ns.jvp(ns.to_device(v)), ns.nv)
    h = 1e-7
    Rp = _residual_at(ns, cs.W0 + h * v)
    Rm = _residual_at(ns, cs.W0 - h * v)
    fd = (Rp - Rm) / (2 * h)
    scale = np.abs(fd).max(axis=0) + 1e-300
    rel = np.abs(Jv - fd) / scale
    print("  per-variable max |Jv - FD| / max|FD|:", " ".join("%.2e" % e for e in rel.max(axis=0)), flush=True)
    return rel


@pytest.mark.parametrize("name,make", [
    ("wedge", lambda: cases.wedge_case(24, 12)),
    ("cylinder_viscous", lambda: cases.cylinder_case(32, 16)),
    ("tetbox_viscous", lambda: cases.tetbox_case(4)),
    ("bump", lambda: cases.bump_case(12, 6, 4)),
    ("plate_prisms", lambda: cases.plate_case(12, 8, 3, first=1e-3)),
])
def test_full_jvp_vs_central_fd_unlimited(name, make):
    """Second-order residual without a limiter: smooth, so the exact tangent
    and central differences (h = 1e-7 of the state) must agree to the FD
    truncation/round-off level."""
    cs = _smooth(make())
    cs.opt.limiter = A.LIMITER_NONE
    rel = _fd_check(cs, "no limiter")
    assert rel.max() < 1e-6


@pytest.mark.parametrize("name,make", [
    ("wedge", lambda: cases.wedge_case(24, 12)),
    ("cylinder_fixed_bcs", lambda: cases.cylinder_case(32, 16)),
    ("tetbox_viscous", lambda: cases.tetbox_case(4)),
    ("plate_prisms", lambda: cases.plate_case(12, 8, 3, first=1e-3)),
])
def test_full_jvp_vs_central_fd_green_gauss(name, make):
    """Green-Gauss gradients in every cell (their tangent is the same face
    sum over the tangent state, BC-fixed face values contributing 0), no
    limiter: the exact tangent against central differences of the GPU
    residual, as for the gradient maps."""
    cs = _smooth(make())
    cs.opt.limiter = A.LIMITER_NONE
    cs.opt.grad_hex = cs.opt.grad_prism = cs.opt.grad_other = A.GRAD_GREENGAUSS
    rel = _fd_check(cs, "green-gauss, no limiter")
    assert rel.max() < 1e-6


@pytest.mark.parametrize("name,make", [
    ("tetbox_vk", lambda: cases.tetbox_case(5)),
    ("bump_vk", lambda: cases.bump_case(12, 6, 4)),
])
def test_full_jvp_vs_central_fd_limited(name, make):
    """With a limiter the residual is only piecewise smooth (min over faces,
    umax/umin switching at local extrema): the tangent takes the active branch
    and central differences straddling a switch differ there by O(1).  The
    Venkatakrishnan function also has a large curvature where dM -> 0 (its
    second derivative grows like 1/dM^2), which costs central differences at
    h = 1e-7 up to ~1e-5 relative on the tet box.  A wrong limiter derivative
    shows as O(1e-2..1) errors on every limited cell, so the bar is: no
    O(1e-3) disagreement on more than 2% of the cells, and a median below
    1e-5.  (The rounding-level check of the limiter tangent is the adjoint
    dot-product test.)

    Minmod on one-cell-thick meshes is not used here: the lsqr z-gradient is
    rounding noise there, whose sign flips phi between 0 and 1 (the
    reference's minmod, ns_limiters.cc:39-43,82-83), a discontinuity rather
    than a kink."""
    cs = _smooth(make())
    rel = _fd_check(cs, "limited")
    per_cell = rel.max(axis=1)
    bad = (per_cell > 1e-3).mean()
    print("  fraction of cells with a limiter switch inside +-h: %.4f, median %.2e" % (bad, np.median(per_cell)), flush=True)
    assert bad < 0.02 and np.median(per_cell) < 1e-5


def test_frozen_jvp_vs_reference_fd_jacobian():
    from oracle.oracle import Oracle
    cs = cases.cylinder_case(32, 16, noise=0.01)
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0, with_jacobian=True)
    JL, JR, _ = o.jacobian()
    m = o.metrics()
    N = cs.n_cells
    v = _direction(N, 9)
    ref = np.zeros((N, 5))
    for f in range(o.info.n_faces):
        p, n = m["face_parent"][f], m["face_neighbor"][f]
        d = JL[f].T @ v[p]
        if m["face_bc"][f] < 0:
            d = d + JR[f].T @ v[n]
            ref[n] -= d
        ref[p] += d
    ns = _ns(cs)
    Jv = ns.to_host(ns.jvp(ns.to_device(v), mode=A.JAC_JVP_FROZEN), 5)
    scale = np.abs(ref).max(axis=0)
    err = (np.abs(Jv - ref) / scale).max()
    print("  frozen JVP vs FD Jacobian matvec: %.2e" % err, flush=True)
    # the reference's absolute FD step (1.5e-7) is ~1e-12 of p ~ 1e5, so the
    # FD Jacobian itself carries ~|F| eps_mach / 1.5e-7 ~ 1e-4 relative noise
    assert err < 1e-3


def test_jvp_linearity():
    cs = cases.tetbox_case(4)
    ns = _ns(cs)
    u = ns.to_device(_direction(cs.n_cells, 1))
    w = ns.to_device(_direction(cs.n_cells, 2))
    Ju = ns.jvp(u.clone()).clone()
    J2u = ns.jvp(2.0 * u).clone()
    Jw = ns.jvp(w.clone()).clone()
    Juw = ns.jvp(u + w).clone()
    N = cs.n_cells
    assert bool((J2u[:, :N] == 2.0 * Ju[:, :N]).all())
    s = (Ju[:, :N].abs() + Jw[:, :N].abs()).max()
    assert float(((Juw - Ju - Jw)[:, :N].abs().max() / s)) < 1e-12


def test_first_order_inviscid_full_equals_frozen():
    cs = cases.wedge_case(16, 8)
    cs.opt.order = A.ORDER_FIRST
    ns = _ns(cs)
    v = ns.to_device(_direction(cs.n_cells, 3))
    a = ns.jvp(v.clone(), mode=A.JAC_JVP_FULL).clone()
    b = ns.jvp(v.clone(), mode=A.JAC_JVP_FROZEN).clone()
    N = cs.n_cells
    assert float((a - b)[:, :N].abs().max() / b[:, :N].abs().max()) < 1e-13


@pytest.mark.parametrize("mode,fp32", [(A.JAC_JVP_FULL, False), (A.JAC_JVP_FULL, True), (A.JAC_JVP_FROZEN, True)])
def test_matrix_free_newton_steps(mode, fp32):
    cs = cases.wedge_case(48, 24)
    ns = _ns(cs)
    opts = A.newton_opts(cfl=10.0, jacobian=mode, fp32=fp32, restart=30, max_iters=30)
    hist = [ns.newton_step(opts).res_norm for _ in range(10)]
    print("  ||R||:", " ".join("%.3e" % h for h in hist), flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] < hist[0]


def _lane_case(name):
    if name == "cylinder":
        return cases.cylinder_case(64, 32)
    if name == "cylinder_bj":
        cs = cases.cylinder_case(64, 32, noise=0.01)
        cs.opt.limiter = A.LIMITER_BJ
        return cs
    if name == "wedge_minmod":
        return cases.wedge_case(48, 24)
    return cases.plate_case(24, 12, 4, noise=0.01, sa=True)


@pytest.mark.parametrize("name", ["cylinder", "plate_sa", "cylinder_bj", "wedge_minmod"])
def test_fp32_tangent_lane_newton_step_matches_fp64_lane(name):
    """The FP32 matrix-free Newton operator carries the tangent's derivative
    lane in FP32 (values stay FP64), and takes the limiters' derivative from
    their linearisation at the step's state (VK on the cylinder and the
    plate, Barth-Jespersen, minmod on the wedge).  One Newton step with it
    against the same step with the FP64 lane (FCFD_TANGENT_FP64=1: the dual-
    number limiter, converted at the operator's boundary), both FP32 GMRES
    to rtol 1e-5: the updates agree to 1e-3 of their size per variable
    (velocity components on the speed's scale), the GMRES iteration counts
    within 2, and the residual after the step within 1e-3 of the residual
    before it (1e-2 on the Mach 2 wedge, whose minmod limiter switches
    branch within the update's rounding and whose GMRES stops at the
    iteration cap: there the FP32 lane with the dual-number limiter lands
    1.2e-2 away, the linearised one 4.5e-3)."""
    import os
    cs = _lane_case(name)
    print("case", cs.name, "N", cs.n_cells, flush=True)
    opts = A.newton_opts(cfl=20.0, jacobian=A.JAC_JVP_FULL, fp32=True, restart=40, max_iters=80, rtol=1e-5)
    out = {}
    for lane in ("fp64", "fp32"):
        if lane == "fp64":
            os.environ["FCFD_TANGENT_FP64"] = "1"
        else:
            os.environ.pop("FCFD_TANGENT_FP64", None)
        ns = _ns(cs)
        st1 = ns.newton_step(opts)
        W1 = ns.state_host(include_ghosts=False)
        ns.assemble_residual()
        R1 = float(np.linalg.norm(ns.residual_host()))
        out[lane] = (W1, st1.krylov_iters, st1.res_norm, R1)
        ns.close()
    os.environ.pop("FCFD_TANGENT_FP64", None)
    d64, d32 = out["fp64"][0] - cs.W0, out["fp32"][0] - cs.W0
    scale = np.abs(d64).max(axis=0)
    scale[1:4] = scale[1:4].max()        # velocity components on the speed's scale (w is rounding noise in 2-D)
    rel = np.abs(d32 - d64).max(axis=0) / (scale + 1e-300)
    print("  per-variable max |dW32 - dW64| / max|dW64|:", " ".join("%.2e" % e for e in rel), flush=True)
    print("  iterations %d / %d, ||R|| after the step %.6e / %.6e" % (out["fp64"][1], out["fp32"][1], out["fp64"][3], out["fp32"][3]),
          flush=True)
    assert rel.max() < 1e-3
    assert abs(out["fp64"][1] - out["fp32"][1]) <= 2
    bar = 1e-2 if name == "wedge_minmod" else 1e-3
    assert abs(out["fp32"][3] - out["fp64"][3]) <= bar * out["fp64"][2]
