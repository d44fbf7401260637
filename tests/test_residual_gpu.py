"""GPU parity of the FP64 residual path against the CPU restatement.

Bar: per cell and variable, |R_gpu - R_oracle| <= 1e-12 * S,
where S = sum over the cell's faces of |face contribution| (the summation
scale; |R| itself cancels to ~0 near steady state).  Ghost states must match
to 1e-14 relative, gradients to 1e-11 of their summation scale.
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

RTOL_R = 1e-12


def _compare(cs, W=None, label=""):
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    W = cs.W0 if W is None else W
    print("case", cs.name, label, "N", cs.n_cells, flush=True)
    o = Oracle(cs.raw, cs.opt)
    Ro, Rabs = o.residual(W)
    Wo, Go, phio = o.fields()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(W)
    ns.assemble_residual()
    ns.status()
    Rg = ns.residual_host()
    Wg = ns.state_host(include_ghosts=True)
    Gg, phig = ns.gradients_limiters_host()
    N = cs.n_cells
    # ghosts
    gscale = np.abs(Wo).max(axis=0)
    assert np.all(np.abs(Wg - Wo) <= 1e-14 * gscale + 1e-300), "ghost state mismatch"
    # gradients: scale = max |G| of the variable over the mesh
    gs = np.abs(Go).max(axis=0) + 1e-300
    gerr = (np.abs(Gg - Go) / gs).max()
    assert gerr <= 1e-11, "gradient mismatch %g" % gerr
    if cs.opt.limiter != A.LIMITER_NONE and cs.opt.order == A.ORDER_SECOND:
        perr = np.abs(phig - phio).max()
        assert perr <= 1e-9, "limiter mismatch %g" % perr
    err = np.abs(Rg - Ro) / (Rabs + 1e-300)
    worst = float(err.max())
    print("  max |dR|/S = %.3e  max|R| = %.3e" % (worst, np.abs(Ro).max()), flush=True)
    assert worst <= RTOL_R, "residual mismatch %g at %s" % (worst, np.unravel_index(err.argmax(), err.shape))
    ns.close()
    o.close()
    return worst


@pytest.mark.parametrize("flux", [A.FLUX_ROE, A.FLUX_AUSM_PLUS_UP, A.FLUX_VAN_LEER, A.FLUX_SD_SLAU])
def test_wedge_fluxes(flux):
    _compare(cases.wedge_case(24, 12, flux=flux), label="flux=%d" % flux)


def test_wedge_full_config():
    _compare(cases.wedge_case(96, 48), label="96x48")


@pytest.mark.parametrize("lim", [A.LIMITER_BJ, A.LIMITER_VK, A.LIMITER_NONE])
def test_wedge_limiters(lim):
    cs = cases.wedge_case(24, 12)
    cs.opt.limiter = lim
    _compare(cs, label="lim=%d" % lim)


def test_first_order():
    cs = cases.wedge_case(24, 12)
    cs.opt.order = A.ORDER_FIRST
    _compare(cs, label="first order")


def test_cylinder_viscous_small():
    _compare(cases.cylinder_case(32, 16, noise=0.01), label="32x16")


def test_cylinder_full_config():
    _compare(cases.cylinder_case(256, 128, noise=0.01), label="256x128")


def test_tetbox_viscous_lsqr():
    _compare(cases.tetbox_case(4), label="n=4")


def test_bump_curvilinear():
    _compare(cases.bump_case(16, 8, 4, noise=0.01), label="16x8x4")


def test_plate_prisms_viscous():
    _compare(cases.plate_case(16, 10, 4, noise=0.01, first=1e-3), label="16x10x4")


def test_pie_collapsed_hexahedra():
    """hexahedra with two repeated nodes at the axis, which the set-up turns
    into prisms (grid_create_elements.cc:274-294, 426-467)"""
    _compare(cases.pie_case(6, 12, 3), label="collapsed hexahedra")


def test_plate_steger_warming():
    cs = cases.plate_case(12, 8, 3, noise=0.01, first=1e-3)
    cs.opt.convective_flux = A.FLUX_SW
    _compare(cs, label="SW")


def _green_gauss_case(name):
    GG = A.GRAD_GREENGAUSS
    if name == "wedge":
        cs = cases.wedge_case(24, 12)
    elif name == "cylinder_fixed_bcs":            # no-slip wall (V fixed), T-fixed inlet, p-fixed outlet
        cs = cases.cylinder_case(32, 16, noise=0.01)
    elif name == "tetbox":
        cs = cases.tetbox_case(4)
    elif name == "plate_prisms_only":             # hexes keep their curvilinear maps
        cs = cases.plate_case(16, 10, 4, noise=0.01, first=1e-3)
        cs.opt.grad_prism = GG
        return cs
    else:
        cs = cases.smooth_state(cases.plate_case(16, 10, 4, sa=True))
    cs.opt.grad_hex, cs.opt.grad_prism, cs.opt.grad_other = GG, GG, GG
    return cs


@pytest.mark.parametrize("name", ["wedge", "cylinder_fixed_bcs", "tetbox", "plate_prisms_only", "plate_sa"])
def test_green_gauss(name):
    """Green-Gauss gradients (the gradients section's ...method=greengauss,
    gradient_maps.cc:45-72; the face sum of Variable::cell_gradient,
    variable.h:193-216) on hexes, tets and prisms, mixed with curvilinear
    maps, with face values fixed by BCs, and with the SA variable: ghosts,
    gradients, limiters and residual against the CPU restatement at the
    residual bar."""
    _compare(_green_gauss_case(name), label="green-gauss " + name)


def test_green_gauss_newton_paths():
    """With Green-Gauss gradients the FD-Jacobian and the matrix-free (exact
    tangent, FP32 Krylov) Newton steps run."""
    from freecfd_amd.ns import NavierStokes
    cs = _green_gauss_case("wedge")
    ns = NavierStokes(cs.raw, cs.opt)
    for opts in (A.newton_opts(cfl=10.0), A.newton_opts(cfl=10.0, jacobian=A.JAC_JVP_FULL, fp32=True)):
        ns.set_state(cs.W0)
        st = ns.newton_step(opts)
        assert np.isfinite(st.res_norm) and st.krylov_iters > 0
    ns.close()


@pytest.mark.parametrize("name,limiter", [
    ("wedge", A.LIMITER_NONE), ("wedge", A.LIMITER_VK),
    pytest.param("wedge", A.LIMITER_MINMOD, marks=pytest.mark.xfail(
        reason="minmod with Green-Gauss gradients on the wedge: <v,Ju> and <J^T v,u> differ by 2.4e-4 "
               "(the same wedge with lsqr maps and minmod agrees to 2e-15); not yet explained")),
    ("cylinder_fixed_bcs", None), ("cylinder_fixed_bcs", A.LIMITER_NONE), ("tetbox", None),
    ("plate_prisms_only", None), ("plate_sa", None), ("plate_sa", A.LIMITER_NONE)])
def test_green_gauss_adjoint_dot_product(name, limiter):
    """The adjoint transposes the Green-Gauss face sum (each face value's cells
    take their weight of n A / V times the gradient adjoint, nothing for the
    variables a BC holds fixed on that face): <v, J u> = <J^T v, u> to 1e-11
    of sum |v J u|, u and v random, J the exact tangent (itself checked
    against central differences with Green-Gauss gradients).  None keeps the
    case's own limiter (Venkatakrishnan)."""
    from freecfd_amd.ns import NavierStokes
    cs = _green_gauss_case(name)
    if limiter is not None:
        cs.opt.limiter = limiter
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    N, nv = cs.n_cells, ns.nv
    rng = np.random.default_rng(5)
    u = rng.normal(size=(N, nv)) * np.array([1e5, 100.0, 100.0, 100.0, 300.0, 1e-4])[:nv]
    v = rng.normal(size=(N, nv))
    Ju = ns.to_host(ns.jvp(ns.to_device(u), mode=A.JAC_JVP_FULL), nv)
    JTv = ns.to_host(ns.vjp(ns.to_device(v)), nv)
    lhs, rhs = float(np.sum(v * Ju)), float(np.sum(JTv * u))
    rel = abs(lhs - rhs) / float(np.sum(np.abs(v * Ju)))
    print("case", name, "<v,Ju> %.15e <J^T v,u> %.15e rel %.2e" % (lhs, rhs, rel), flush=True)
    assert rel < 1e-11
    ns.close()


def _unpack_pairs(T, comps, ld, perm, rows):
    """Row-pair device array -> [rows][comps] in reference numbering."""
    a = T.cpu().numpy().reshape(-1, ld, 2)
    return np.stack([a[c // 2, perm[:rows], c % 2] for c in range(comps)], axis=1)


@pytest.mark.parametrize("name", ["wedge_vk", "first_order", "tetbox_lsqr", "plate_sa"])
def test_gradients_limiters_alone(name):
    """fcfd_gradients_limiters (ghosts, gradients and limiters of W, into the
    caller's buffers) against the CPU restatement at the residual test's
    tolerances, and bitwise against what the full residual leaves in the
    context's buffers."""
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    if name == "wedge_vk":
        cs = cases.wedge_case(24, 12)
        cs.opt.limiter = A.LIMITER_VK
    elif name == "first_order":
        cs = cases.wedge_case(24, 12)
        cs.opt.order = A.ORDER_FIRST
    elif name == "tetbox_lsqr":
        cs = cases.tetbox_case(4)
    else:
        cs = cases.smooth_state(cases.plate_case(16, 10, 4, sa=True))
    print("case", cs.name, name, flush=True)
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0)
    Wo, Go, phio = o.fields()
    o.close()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    G, PHI = ns.gradients_limiters()
    ns.status()
    N, B, nv, ld = cs.n_cells, ns.info.n_ghosts, ns.nv, ns.info.ld_cell
    perm = ns.metrics()["cell_perm"]
    Gg = _unpack_pairs(G, 3 * nv, ld, perm, N + B)
    Wg = ns.state_host(include_ghosts=True)
    assert np.all(np.abs(Wg - Wo) <= 1e-14 * np.abs(Wo).max(axis=0) + 1e-300), "ghost state mismatch"
    gerr = (np.abs(Gg - Go) / (np.abs(Go).max(axis=0) + 1e-300)).max()
    limited = cs.opt.limiter != A.LIMITER_NONE and cs.opt.order == A.ORDER_SECOND
    perr = 0.0
    if limited:
        phig = _unpack_pairs(PHI, nv, ld, perm, N)
        perr = np.abs(phig - phio).max()
    print("  gradients %.3e  limiters %.3e" % (gerr, perr), flush=True)
    assert gerr <= 1e-11
    assert perr <= 1e-9
    # the residual computes the same fields into the context's buffers
    ns.assemble_residual()
    Gr, phir = ns.gradients_limiters_host()
    assert np.array_equal(Gr, Gg)
    if limited:
        assert np.array_equal(phir, phig)
    ns.close()
