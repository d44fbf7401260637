"""Adjoint sensitivities of a wall drag coefficient to freestream parameters.

For a converged steady state W(alpha) of R(W, alpha) = 0 and an output
J(W, alpha), the total derivative is

    dJ/dalpha = dJ/dalpha|_W + lam . dR/dalpha|_W,   (dR/dW)^T lam = -dJ/dW,

one adjoint solve for any number of parameters.  Here J is the drag (skin
friction on a no-slip wall, whose normal is orthogonal to the flow) in
coefficient form,

    C_D = F_x / (q_inf S_wall),   q_inf = 1/2 rho_inf U_inf^2,

with F_x the x force on the wall BC from the face loop (loads.h:33-39,
ns_assemble_linear_system.cc:109-119).  dJ/dW comes from the exact adjoint
kernels (fcfd_vjp with a load weight), lam from fcfd_adjoint_step.  R and
J depend on a parameter only through BC values (and q_inf); the two partials
at fixed W are exact, from the tangent kernels with the BC values as dual
numbers (fcfd_bc_tangent: dR/dalpha and dF_x/dalpha through the ghosts,
gradients, limiters and boundary fluxes), plus the closed form of
dq_inf/dalpha.

Parameters are inlet BC values: the Mach number (inlet velocity = M a_inf
along x at fixed T_inf, so q_inf moves with it) and the inlet nu~ of the SA
model.  The reference has no adjoint; this is a build-only capability over
its residual path.
"""
import numpy as np

from . import _abi as A


class InletMach:
    """alpha = M_inf: inlet velocity M a_inf e_x on BC `inlet`."""

    name = "mach"

    def __init__(self, inlet, a_inf, rho_inf):
        self.inlet, self.a_inf, self.rho_inf = inlet, a_inf, rho_inf

    def apply(self, ns, value):
        ns.set_bc(self.inlet, V=(value * self.a_inf, 0.0, 0.0))

    def q_inf(self, value, base_q):
        u = value * self.a_inf
        return 0.5 * self.rho_inf * u * u

    def direction(self):
        """(BC index, d(BC values)/d alpha)"""
        return self.inlet, {"V": (self.a_inf, 0.0, 0.0)}

    def dq_inf(self, value, base_q):
        return self.rho_inf * self.a_inf * self.a_inf * value


class InletNut:
    """alpha = nu~ at the inlet (SA model)."""

    name = "nut"

    def __init__(self, inlet):
        self.inlet = inlet

    def apply(self, ns, value):
        ns.set_bc(self.inlet, nut=value)

    def q_inf(self, value, base_q):
        return base_q

    def direction(self):
        return self.inlet, {"nut": 1.0}

    def dq_inf(self, value, base_q):
        return 0.0


def wall_area(ns, wall):
    m = ns.metrics()
    return float(np.sum(m["face_area"][m["face_bc"] == wall]))


def drag_coefficient(ns, wall, q_inf, area, W=None):
    return float(ns.loads(W)[wall, 0]) / (q_inf * area)


def drag_weight(n_bc, wall, q_inf, area):
    """fx_weight of fcfd_vjp for d C_D / dW: the force is -sum (D - C)[1..3]."""
    w = np.zeros((n_bc, 5))
    w[wall, 1] = -1.0 / (q_inf * area)
    return w


def partials(ns, wall, prm, val, q_inf, area, lam=None):
    """The exact partials at the fixed state ns.W: (lam . dR/dalpha or the
    dR/dalpha tensor when lam is None, dC_D/dalpha|_W)"""
    index, d = prm.direction()
    dR, dl = ns.bc_tangent(index, **d)
    n = ns.info.n_cells
    J = drag_coefficient(ns, wall, q_inf, area)
    explicit = float(dl[wall, 0]) / (q_inf * area) - J * prm.dq_inf(val, q_inf) / q_inf
    if lam is None:
        return dR, explicit
    return float(ns.torch.sum(lam[:, :n] * dR[:, :n])), explicit


def adjoint_sensitivities(ns, wall, params, values, q_inf, adjoint_opts=None, adjoint_steps=200,
                          adjoint_drop=1e-11, cfl_initial=10.0, cfl_growth=2.0, cfl_target=1e6):
    """d C_D / d alpha for each parameter (objects with apply/q_inf/
    direction/dq_inf, current values `values`) at the converged state held
    in ns.W.  Returns a dict with C_D, the adjoint solve's history and, per
    parameter, the total derivative and its two parts."""
    torch = ns.torch
    area = wall_area(ns, wall)
    n_bc = int(ns._opt.n_bc)
    J0 = drag_coefficient(ns, wall, q_inf, area)
    zero = torch.zeros_like(ns.W)
    g = ns.vjp(zero, fx_weight=drag_weight(n_bc, wall, q_inf, area)).clone()
    opts = adjoint_opts or A.newton_opts(cfl=cfl_initial, jacobian=A.JAC_FD, fp32=False, restart=30, max_iters=60,
                                         rtol=1e-6, precond=A.PC_BLOCK_JACOBI)
    lam, hist = ns.adjoint_solve(g, opts, max_steps=adjoint_steps, drop=adjoint_drop, cfl_initial=cfl_initial,
                                 cfl_growth=cfl_growth, cfl_target=cfl_target)
    out = {"C_D": J0, "area": area, "adjoint": hist, "params": {}}
    for prm, val in zip(params, values):
        lam_term, dJda = partials(ns, wall, prm, val, q_inf, area, lam)
        out["params"][prm.name] = {"value": val, "total": dJda + lam_term, "explicit": dJda, "adjoint_term": lam_term}
    ns.assemble_residual()
    return out


def finite_difference(ns, wall, prm, val, q_inf, h, solve, area=None):
    """Central difference of C_D in one parameter with the flow re-converged
    at val + h and val - h by `solve(ns)` (from the current state); ns.W is
    restored to the state it held on entry."""
    area = wall_area(ns, wall) if area is None else area
    W0 = ns.W.clone()
    res = []
    for v in (val + h, val - h):
        ns.W.copy_(W0)
        prm.apply(ns, v)
        info = solve(ns)
        res.append((drag_coefficient(ns, wall, prm.q_inf(v, q_inf), area), info))
    prm.apply(ns, val)
    ns.W.copy_(W0)
    ns.torch.cuda.synchronize()
    return (res[0][0] - res[1][0]) / (2.0 * h), res
