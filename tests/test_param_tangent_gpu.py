"""Exact derivatives of the residual and the loads with respect to BC values
(fcfd_bc_tangent, fcfd_vjp_params) against central differences of the GPU
residual and loads with the BC value moved by fcfd_set_bc.

The channel variants run every inlet, outlet and wall kind (velocity and
mass-flow inlets with each specification, the stagnation inlet, p/T/rho and
reverse-flow outlets, adiabatic, FIXED_T, FIXED_Q and slip walls), with SA
for the inlet nu~.  Each BC value that enters its BC kind is differentiated
(the others must give exactly 0).  Bar, per variable: max |d_exact - d_fd|
<= 1e-4 of the largest |d_fd| of that variable plus the central
difference's rounding, 64 eps max|R_v| / h (relative step 1e-6); the same
for the loads.  (The VK limiter's min/max switches make the residual only
piecewise smooth in a BC value; a missing path gives O(1) errors.)  The
parameter VJP must equal <Rbar, dR/dv> of the tangent, and its load-weight
part the loads' derivative.
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

VARIANTS = sorted(cases.BC_VARIANTS)
FIELDS = ("p", "T", "rho", "V", "p_total", "T_total", "mdot", "qdot", "nut")


def _value(bc, field):
    return float(bc.V[0]) if field == "V" else float(getattr(bc, field))


def _set(ns, b, field, value, base):
    if field == "V":
        ns.set_bc(b, V=(value, base.V[1], base.V[2]))
    else:
        ns.set_bc(b, **{field: value})


def _fd(ns, b, field, h):
    base = A.BC.from_buffer_copy(ns._opt.bc[b])
    v0 = _value(base, field)
    out = []
    for v in (v0 + h, v0 - h):
        _set(ns, b, field, v, base)
        R = ns.assemble_residual().clone()
        out.append((R, ns.loads()))
    _set(ns, b, field, v0, base)
    dR = (out[0][0] - out[1][0]) / (2 * h)
    dl = (out[0][1] - out[1][1]) / (2 * h)
    return dR, dl


@pytest.mark.parametrize("variant", VARIANTS)
def test_bc_tangent_vs_central_differences(variant):
    _check_bc_tangent(cases.bc_variant_case(variant, sa=True))


@pytest.mark.parametrize("variant", ["vel_p__out_t__wall_fixed_t", "stagnation__out_t_no_reverse__wall_fixed_t",
                                     "mdot_state__out_p__wall_adiabatic"])
def test_bc_tangent_green_gauss_vs_central_differences(variant):
    """Green-Gauss gradients in every cell: a face value a BC holds fixed
    (p, V or T) is that BC's value, so its tangent is the seeded direction.
    Without a limiter the residual is smooth in the BC values, so central
    differences meet the bar with room (a missing path gives O(1) errors);
    with Venkatakrishnan's the FIXED_T wall's T derivative sits at 1e-4."""
    cs = cases.bc_variant_case(variant, sa=True)
    cs.opt.grad_hex = cs.opt.grad_prism = cs.opt.grad_other = A.GRAD_GREENGAUSS
    cs.opt.limiter = A.LIMITER_NONE
    _check_bc_tangent(cs)


def _check_bc_tangent(cs):
    from freecfd_amd.ns import NavierStokes
    print("case", cs.name, flush=True)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    n = cs.n_cells
    Rmax = np.abs(ns.assemble_residual()[:, :n].cpu().numpy()).max(axis=1)
    checked = 0
    for b in range(3):   # inlet, outlet, wall
        for field in FIELDS:
            v0 = _value(cs.opt.bc[b], field)
            h = 1e-6 * abs(v0) if v0 != 0.0 else 1e-6
            d = {"V": (1.0, 0.0, 0.0)} if field == "V" else {field: 1.0}
            dR, dl = ns.bc_tangent(b, **d)
            dR = dR[:, :n].cpu().numpy()
            fR, fl = _fd(ns, b, field, h)
            fR = fR[:, :n].cpu().numpy()
            scale = np.abs(fR).max(axis=1)
            err = np.abs(dR - fR).max(axis=1)
            if scale.max() == 0.0:
                assert np.abs(dR).max() == 0.0, (b, field)
                continue
            # per variable: 1e-4 of its largest derivative, plus the central
            # difference's rounding, 64 eps max|R_v| / h
            noise = 64 * np.finfo(float).eps * Rmax / h
            rel = float((err / (1e-4 * scale + noise + 1e-300)).max()) * 1e-4
            # the loads' central difference carries rounding of about
            # eps |loads| / h: the bar is 1e-4 of |dloads| plus 64 times that
            lnoise = 64 * np.finfo(float).eps * np.abs(ns.loads()).max() / h
            lerr = float(np.abs(dl - fl).max())
            lbar = 1e-4 * np.abs(fl).max() + lnoise
            print("  BC %d %-7s |dR| %.3e rel %.2e; loads |d| %.3e err %.2e (bar %.2e)" % (b, field, scale.max(), rel,
                                                                                        np.abs(fl).max(), lerr, lbar), flush=True)
            assert rel <= 1e-4, (b, field)
            assert lerr <= lbar, (b, field)
            checked += 1
    assert checked >= 3


def test_vjp_params_equals_the_tangent_and_the_loads_weight():
    from freecfd_amd.ns import NavierStokes
    torch = pytest.importorskip("torch")
    cs = cases.bc_variant_case(VARIANTS[0], sa=True)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    n = cs.n_cells
    rng = np.random.default_rng(4)
    Rbar = torch.zeros_like(ns.W)
    Rbar[:, :n] = torch.as_tensor(rng.normal(size=(ns.nv, n)), device=ns.W.device)
    pbar = ns.vjp_params(Rbar)
    fields = ["p", "T", "rho", ("V", 0), ("V", 1), ("V", 2), "p_total", "T_total", "mdot", "qdot", "nut"]
    for b in range(cs.opt.n_bc):
        for q, fld in enumerate(fields):
            if isinstance(fld, tuple):
                v = [0.0, 0.0, 0.0]
                v[fld[1]] = 1.0
                dR, _ = ns.bc_tangent(b, V=tuple(v))
            else:
                dR, _ = ns.bc_tangent(b, **{fld: 1.0})
            want = float(torch.sum(Rbar[:, :n] * dR[:, :n]))
            assert pbar[b, q] == pytest.approx(want, rel=1e-12, abs=1e-12 * (abs(want) + 1e-30))
    # the load-weight part: Rbar = 0, fx_weight = the wall's x-force weight
    w = np.zeros((cs.opt.n_bc, 5))
    w[2, 1] = -1.0
    pw = ns.vjp_params(torch.zeros_like(ns.W), fx_weight=w)
    for b in range(3):
        _, dl = ns.bc_tangent(b, p=1.0)
        assert pw[b, 0] == pytest.approx(dl[2, 0], rel=1e-12, abs=1e-300)
