"""Adjoint sensitivities of the wall drag coefficient of the SA flat plate to
the inlet Mach number and the inlet nu~, against central differences of
re-converged flows.

The flow (SA-RANS plate 16x12x2, Mach 0.3, Re_L 5e6, AUSM+up, unlimited second
order) is converged with the exact-tangent Newton solver; the adjoint
(dR/dW)^T lam = -dC_D/dW is solved with fcfd_adjoint_step (exact transpose
operator, transposed block-Jacobi); dC_D/dalpha = dC_D/dalpha|_W +
lam . dR/dalpha|_W, with both partials exact (fcfd_bc_tangent: the inlet's
BC values as dual numbers through ghosts, gradients, limiters and boundary
fluxes).  The check: re-converge the flow at alpha (1 +- 1e-3) and
difference C_D.  Agreement is limited by how far the primal converges (its
residual stalls near 3e-9 absolute from 2.5 at the start of the transient),
which bounds the difference quotient's error at about 1e-5 relative.

Roe on this plate does not reach a steady state in pseudo-time (it diverges
once CFL exceeds a few hundred, and at CFL 5 when the top layers are prisms
with curvilinear gradients), so the sensitivity case uses AUSM+up.
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases
from freecfd_amd import sensitivity as S

pytestmark = pytest.mark.gpu

WALL, INLET = 2, 0


def _popts(cfl):
    return A.newton_opts(cfl=cfl, jacobian=A.JAC_JVP_FULL, fp32=False, restart=50, max_iters=200, rtol=1e-10,
                         precond=A.PC_BLOCK_SGS, sweeps=2)


@pytest.fixture(scope="module")
def converged():
    from freecfd_amd.ns import NavierStokes
    cs = cases.plate_case(16, 12, 2, first=1e-2, sa=True)
    cs.opt.convective_flux = A.FLUX_AUSM_PLUS_UP
    cs.opt.limiter = A.LIMITER_NONE
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    r = ns.solve(_popts(1.0), max_steps=70, drop=1e-14, cfl_initial=1.0, cfl_growth=1.2, cfl_target=1e8)
    h = r["res_norm"]
    print("  primal: %d steps, ||R|| %.3e -> %.3e (max %.3e)" % (r["steps"], h[0], h[-1], max(h)), flush=True)
    assert h[-1] < 2e-8 and h[-1] < 1e-8 * max(h)
    m = cs.meta
    rho = m["p"] / (m["R"] * m["T"])
    a = m["U"] / m["mach"]
    q = 0.5 * rho * m["U"] ** 2
    return ns, cs, q, [S.InletMach(INLET, a, rho), S.InletNut(INLET)], [m["mach"], m["nut"]]


def test_adjoint_converges_and_matches_fd(converged):
    ns, cs, q, params, vals = converged
    out = S.adjoint_sensitivities(ns, WALL, params, vals, q, adjoint_steps=40, adjoint_drop=1e-5)
    ah = out["adjoint"]
    print("  C_D %.10e, adjoint %s in %d steps, ||rho||/||g|| %.2e" % (out["C_D"], ah["status"], ah["steps"], ah["drop"]),
          flush=True)
    assert ah["status"] == "converged"
    assert out["C_D"] > 0.0

    def resolve(n):
        rr = n.solve(_popts(1e8), max_steps=12, drop=1e-30)
        return rr["res_norm"][-1]

    for prm, val in zip(params, vals):
        d = out["params"][prm.name]
        fd, res = S.finite_difference(ns, WALL, prm, val, q, 1e-3 * val, resolve)
        rel = abs(fd - d["total"]) / abs(fd)
        print("  d C_D / d %s: adjoint %.10e (explicit %.3e + adjoint term %.3e), central FD %.10e, rel %.2e, "
              "re-solved ||R|| %s" % (prm.name, d["total"], d["explicit"], d["adjoint_term"], fd, rel,
                                      ", ".join("%.1e" % x[1] for x in res)), flush=True)
        assert all(x[1] < 2e-8 for x in res)
        assert rel < 1e-4


def test_adjoint_step_block_sgs_and_multigrid_match_block_jacobi(converged):
    """One adjoint step from lam = 0 preconditioned by the transposed block SGS
    sweep, and by the transposed multigrid V-cycle, reaches the same update as
    with the transposed block Jacobi (the preconditioner changes the Krylov
    iterations, not the step's linear system), in fewer iterations: a sweep or
    cycle over the wrong transpose would still converge, but no faster than
    block Jacobi."""
    ns, cs, q, params, vals = converged
    torch = ns.torch
    area = S.wall_area(ns, WALL)
    g = ns.vjp(torch.zeros_like(ns.W), fx_weight=S.drag_weight(int(ns._opt.n_bc), WALL, q, area)).clone()
    out = {}
    for pc in (A.PC_BLOCK_JACOBI, A.PC_BLOCK_SGS, A.PC_MULTIGRID):
        lam = torch.zeros_like(ns.W)
        opts = A.newton_opts(cfl=20.0, jacobian=A.JAC_FD, fp32=False, restart=120, max_iters=480, rtol=1e-12, abstol=0.0,
                             knoll=False, precond=pc, sweeps=1)
        st = ns.adjoint_step(g, lam, opts)
        out[pc] = (lam.clone(), st.krylov_iters, st.krylov_converged)
        print("  precond %d: %d Krylov iterations, converged %d, rnorm %.3e/%.3e" % (pc, st.krylov_iters, st.krylov_converged,
                                                                                 st.krylov_rnorm, st.krylov_r0norm), flush=True)
    lj, ij, cj = out[A.PC_BLOCK_JACOBI]
    ls, isgs, csgs = out[A.PC_BLOCK_SGS]
    lm, img, cmg = out[A.PC_MULTIGRID]
    err = float((ls - lj).abs().max() / lj.abs().max())
    errm = float((lm - lj).abs().max() / lj.abs().max())
    print("  max |lam_SGS - lam_BJ| / max |lam_BJ| = %.3e, multigrid %.3e" % (err, errm), flush=True)
    assert cj and csgs and cmg
    assert err < 1e-6 and errm < 1e-6
    assert isgs < ij and img < ij


def test_set_bc_roundtrip(converged):
    """fcfd_set_bc changes the residual through the BC table and restores it
    exactly when the old values come back."""
    ns, cs, q, params, vals = converged
    R0 = ns.assemble_residual().clone()
    V0 = tuple(cs.opt.bc[INLET].V)
    ns.set_bc(INLET, V=(V0[0] * 1.01, V0[1], V0[2]))
    R1 = ns.assemble_residual().clone()
    ns.set_bc(INLET, V=V0)
    R2 = ns.assemble_residual().clone()
    assert float((R1 - R0).abs().max()) > 0.0
    assert bool((R2 == R0).all())
