"""GPU parity for every boundary-condition kind against the CPU restatement.

Each variant of cases.BC_VARIANTS puts a different inlet, outlet and wall
on a small jittered viscous channel whose sheared inflow makes some inlet
and outlet faces see reverse flow:
  - velocity inlets with p, T or rho given; mass-flow inlets with p, T, rho
    or the full state; the stagnation inlet;
  - outlets with p, T, rho or nothing given, with the DAMP_REVERSE,
    NO_REVERSE and FORCE_SUPERSONIC kinds;
  - no-slip walls (adiabatic, FIXED_T, FIXED_Q) and a slip wall.
Every branch of apply_bcs (ns_apply_bcs.cc:25-258) and of the FIXED_Q heat
flux (ns_diffusive_face_flux.cc:58-60) therefore runs on the GPU here.  The
checks: the residual against the oracle (<= 1e-12 of the summed
contributions), the FD Jacobian against the oracle (1e-6 per block: the BCs'
entropy extrapolations use pow, see test_jacobian_gpu.py), the surface loads
against the oracle, the exact tangent against central differences of the
GPU residual, and the dot-product test of tangent and adjoint.  A mass-flow
inlet turned supersonic raises the BC flag on the GPU where the oracle
stops (ns_apply_bcs.cc:99-104).
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

VARIANTS = sorted(cases.BC_VARIANTS)
SCALES = np.array([1e5, 100.0, 100.0, 100.0, 300.0, 1e-4])


def _ns(cs):
    from freecfd_amd.ns import NavierStokes
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    return ns


@pytest.mark.parametrize("sa", [False, True])
@pytest.mark.parametrize("variant", VARIANTS)
def test_bc_residual_parity(variant, sa):
    from oracle.oracle import Oracle
    cs = cases.bc_variant_case(variant, sa=sa)
    print("case", cs.name, flush=True)
    Ro, Rabs = Oracle(cs.raw, cs.opt).residual(cs.W0)
    ns = _ns(cs)
    ns.assemble_residual()
    ns.status()
    Rg = ns.residual_host()
    err = float((np.abs(Rg - Ro) / (Rabs + 1e-300)).max())
    print("  max |R_gpu - R_oracle| / sum|contrib| = %.2e" % err, flush=True)
    assert err <= 1e-12


def _blockwise_err(a, b):
    keep = np.abs(b).reshape(len(b), -1).max(axis=1) > 0
    a, b = a[keep], b[keep]
    scale = np.abs(b).reshape(len(b), -1).max(axis=1)
    return float((np.abs(a - b).reshape(len(b), -1).max(axis=1) / scale).max())


@pytest.mark.parametrize("variant", VARIANTS)
def test_bc_fd_jacobian_parity(variant):
    from oracle.oracle import Oracle
    cs = cases.bc_variant_case(variant)
    print("case", cs.name, flush=True)
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0, with_jacobian=True)
    JLo, JRo, Do = o.jacobian()
    ns = _ns(cs)
    ns.fd_jacobian()
    ns.status()
    JLg, JRg, Dg = ns.jacobian_host()
    e = (_blockwise_err(JLg, JLo), _blockwise_err(JRg, JRo), _blockwise_err(Dg, Do))
    print("  JL %.2e  JR %.2e  D %.2e" % e, flush=True)
    assert max(e) <= 1e-6


@pytest.mark.parametrize("variant", VARIANTS)
def test_bc_loads_parity(variant):
    from oracle.oracle import Oracle
    cs = cases.bc_variant_case(variant)
    print("case", cs.name, flush=True)
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0)
    ref = o.loads()
    got = _ns(cs).loads()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("  loads err %.2e" % err, flush=True)
    assert np.abs(ref[2]).max() > 0 and err <= 1e-13


@pytest.mark.parametrize("variant", VARIANTS)
def test_bc_tangent_vs_central_fd(variant):
    """Unlimited second order: the exact tangent through every BC branch
    against central differences (h = 1e-7 of the state scales)."""
    cs = cases.bc_variant_case(variant, limiter=A.LIMITER_NONE)
    print("case", cs.name, flush=True)
    ns = _ns(cs)
    rng = np.random.default_rng(4)
    v = rng.normal(size=cs.W0.shape) * SCALES[:ns.nv]
    Jv = ns.to_host(ns.jvp(ns.to_device(v)), ns.nv)
    h = 1e-7
    def R_at(W):
        ns.set_state(W)
        ns.assemble_residual()
        return ns.residual_host()
    fd = (R_at(cs.W0 + h * v) - R_at(cs.W0 - h * v)) / (2 * h)
    rel = np.abs(Jv - fd).max(axis=0) / (np.abs(fd).max(axis=0) + 1e-300)
    print("  per-variable max |Jv - FD| / max|FD|:", " ".join("%.2e" % e for e in rel), flush=True)
    assert rel.max() < 1e-6


@pytest.mark.parametrize("sa", [False, True])
@pytest.mark.parametrize("variant", VARIANTS)
def test_bc_dot_product(variant, sa):
    cs = cases.bc_variant_case(variant, sa=sa)
    print("case", cs.name, flush=True)
    ns = _ns(cs)
    N, nv = cs.n_cells, ns.nv
    rng = np.random.default_rng(11)
    u = rng.normal(size=(N, nv)) * SCALES[:nv]
    w = rng.normal(size=(N, nv))
    Ju = ns.to_host(ns.jvp(ns.to_device(u), mode=A.JAC_JVP_FULL), nv)
    JTw = ns.to_host(ns.vjp(ns.to_device(w)), nv)
    lhs, rhs = float(np.sum(w * Ju)), float(np.sum(JTw * u))
    rel = abs(lhs - rhs) / float(np.sum(np.abs(w * Ju)))
    print("  <w,Ju> %.15e  <J^T w,u> %.15e  rel %.2e" % (lhs, rhs, rel), flush=True)
    ns.status()
    assert rel <= 1e-11


@pytest.mark.parametrize("spec", [A.SPEC_P, A.SPEC_STATE])
def test_supersonic_mass_flow_inlet_raises_the_bc_flag(spec):
    """ns_apply_bcs.cc:99-104 stops for every specification: the GPU raises
    FCFD_EBC through fcfd_status, the oracle refuses the state."""
    from oracle.oracle import Oracle
    cs = cases.bc_variant_case("mdot_state__out_p__wall_adiabatic" if spec == A.SPEC_STATE else
                               "mdot_p__out_no_reverse__wall_fixed_t")
    print("case", cs.name, "supersonic inflow", flush=True)
    a = np.sqrt(1.4 * cs.meta["R"] * cs.meta["T"])
    W = cs.W0.copy()
    W[:, 1] = 1.3 * a                           # every inlet face: Mach_L = -u_n / a = 1.3
    with pytest.raises(RuntimeError, match="oracle error -3"):
        Oracle(cs.raw, cs.opt).residual(W)
    ns = _ns(cs)
    ns.set_state(W)
    ns.assemble_residual()
    with pytest.raises(RuntimeError, match="libfcfd error -3"):
        ns.status()
