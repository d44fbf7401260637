"""Face-average stencils wider than a face's two cells (WTLI with tetrahedron
and triangle fits, IDW) on the device: the viscous face gradients average
the stencil's cell gradients in the reference's key order
(Variable<Vec3D>::face_calculate, variable.h:134-147).

- residual against the oracle at 1e-12 of the per-cell sum of |contributions|
  (and the set-up's stencils bitwise, through the same metrics), on a 3-D tet
  box (WTLI stencil 8: tetrahedron fits), the 2-D cylinder (WTLI 6: triangle
  fits) and the SA plate (IDW 8, hexes and prisms);
- the FD Jacobian against the oracle (1e-6 per block: Sutherland's pow);
- the exact tangent against central differences of the GPU residual;
- the adjoint (dot-product test against the tangent).
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

RTOL_R = 1e-12


def _wide(cs, method, stencil):
    cs.opt.interp_method = method
    cs.opt.stencil_size = stencil
    return cs


CASES = [
    ("tetbox_wtli8", lambda: _wide(cases.tetbox_case(4), A.INTERP_WTLI, 8)),
    ("cylinder_wtli6", lambda: _wide(cases.cylinder_case(24, 12, noise=0.01), A.INTERP_WTLI, 6)),
    ("plate_sa_idw8", lambda: _wide(cases.plate_case(10, 6, 3, noise=0.01, sa=True, first=1e-3), A.INTERP_IDW, 8)),
]


@pytest.mark.parametrize("name,make", CASES)
def test_residual_with_wide_stencils_matches_the_oracle(name, make):
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = make()
    print("case", name, "N", cs.n_cells, flush=True)
    o = Oracle(cs.raw, cs.opt)
    m = o.metrics()
    widths = np.diff(m["avg_ptr"])
    assert widths.max() > 2, "the case must have stencils wider than two cells"
    Ro, Rabs = o.residual(cs.W0)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    ns.assemble_residual()
    ns.status()
    Rg = ns.residual_host()
    err = np.abs(Rg - Ro) / (Rabs + 1e-300)
    print("  widest stencil %d, max |dR|/S = %.3e" % (widths.max(), err.max()), flush=True)
    assert err.max() <= RTOL_R
    ns.close()
    o.close()


@pytest.mark.parametrize("name,make", CASES)
def test_fd_jacobian_with_wide_stencils_matches_the_oracle(name, make):
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = make()
    print("case", name, flush=True)
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0, with_jacobian=True)
    JLo, JRo, Do = o.jacobian()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    ns.fd_jacobian()
    JLg, JRg, Dg = ns.jacobian_host()

    def blockwise(a, b):
        keep = np.abs(b).reshape(len(b), -1).max(axis=1) > 0
        a, b = a[keep], b[keep]
        scale = np.abs(b).reshape(len(b), -1).max(axis=1)
        return (np.abs(a - b).reshape(len(b), -1).max(axis=1) / scale).max()

    e = (blockwise(JLg, JLo), blockwise(JRg, JRo), blockwise(Dg, Do))
    print("  JL %.2e  JR %.2e  D %.2e" % e, flush=True)
    assert max(e) <= 1e-6
    ns.close()
    o.close()


@pytest.mark.parametrize("name,make", CASES[:2])
def test_tangent_with_wide_stencils_vs_central_differences(name, make):
    from freecfd_amd.ns import NavierStokes
    cs = cases.smooth_state(make(), 21, 0.02)
    cs.opt.limiter = A.LIMITER_NONE
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    rng = np.random.default_rng(4)
    v = rng.normal(size=(cs.n_cells, ns.nv)) * np.array([1e5, 100.0, 100.0, 100.0, 300.0, 1e-4])[:ns.nv]
    Jv = ns.to_host(ns.jvp(ns.to_device(v)), ns.nv)
    h = 1e-7
    R = []
    for sgn in (1, -1):
        ns.set_state(cs.W0 + sgn * h * v)
        ns.assemble_residual()
        R.append(ns.residual_host())
    fd = (R[0] - R[1]) / (2 * h)
    rel = np.abs(Jv - fd) / (np.abs(fd).max(axis=0) + 1e-300)
    print("case", name, "max |Jv - FD| / max|FD| per variable:", " ".join("%.2e" % x for x in rel.max(axis=0)), flush=True)
    assert rel.max() < 1e-6
    ns.close()


@pytest.mark.parametrize("name,make", CASES)
def test_adjoint_with_wide_stencils_dot_product(name, make):
    """The adjoint transposes the wide face averages (each stencil cell takes
    its weight of the face-gradient adjoint): <v, J u> = <J^T v, u> to 1e-11
    of sum |v J u|, with u, v random and J the exact tangent."""
    from freecfd_amd.ns import NavierStokes
    cs = make()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    N, nv = cs.n_cells, ns.nv
    rng = np.random.default_rng(11)
    u = rng.normal(size=(N, nv)) * np.array([1e5, 100.0, 100.0, 100.0, 300.0, 1e-4])[:nv]
    v = rng.normal(size=(N, nv))
    Ju = ns.to_host(ns.jvp(ns.to_device(u), mode=A.JAC_JVP_FULL), nv)
    JTv = ns.to_host(ns.vjp(ns.to_device(v)), nv)
    lhs, rhs = float(np.sum(v * Ju)), float(np.sum(JTv * u))
    rel = abs(lhs - rhs) / float(np.sum(np.abs(v * Ju)))
    print("case", name, "<v,Ju> %.15e <J^T v,u> %.15e rel %.2e" % (lhs, rhs, rel), flush=True)
    assert rel < 1e-11
    ns.close()


def test_fp32_tangent_lane_with_wide_stencils_matches_fp64_lane():
    """The FP32 matrix-free Newton operator with the wide stencils' face
    gradients in its FP32 lane: one Newton step against the same step with
    the FP64 lane (FCFD_TANGENT_FP64=1), FP32 GMRES to rtol 1e-5; the updates
    agree to 1e-3 of their size per variable (velocity on the speed's scale)
    and the GMRES iteration counts within 2."""
    import os
    from freecfd_amd.ns import NavierStokes
    cs = CASES[1][1]()
    opts = A.newton_opts(cfl=20.0, jacobian=A.JAC_JVP_FULL, fp32=True, restart=40, max_iters=80, rtol=1e-5)
    out = {}
    for lane in ("fp64", "fp32"):
        if lane == "fp64":
            os.environ["FCFD_TANGENT_FP64"] = "1"
        else:
            os.environ.pop("FCFD_TANGENT_FP64", None)
        ns = NavierStokes(cs.raw, cs.opt)
        ns.set_state(cs.W0)
        st = ns.newton_step(opts)
        out[lane] = (ns.state_host(include_ghosts=False), st.krylov_iters)
        ns.close()
    os.environ.pop("FCFD_TANGENT_FP64", None)
    d64, d32 = out["fp64"][0] - cs.W0, out["fp32"][0] - cs.W0
    scale = np.abs(d64).max(axis=0)
    scale[1:4] = scale[1:4].max()
    rel = np.abs(d32 - d64).max(axis=0) / (scale + 1e-300)
    print("  per-variable max |dW32 - dW64| / max|dW64|:", " ".join("%.2e" % e for e in rel), flush=True)
    assert rel.max() < 1e-3
    assert abs(out["fp64"][1] - out["fp32"][1]) <= 2
