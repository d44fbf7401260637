"""The five benchmark configurations as (raw mesh, options, initial state).

Inputs mirror what a Free CFD input deck sets (src/read_inputs.cc defaults,
src/navier_stokes/ns_set_bcs.cc for BC values); every state array is N x 5
primitive (p, u, v, w, T) in raw (= reference) cell order.
"""
import numpy as np

from . import _abi as A
from . import meshgen


def gas_R(Mw):
    return A.UNIV_GAS_CONST / Mw


# ---------------------------------------------------------------------------
# boundary conditions (what NavierStokes::set_bcs derives from an input deck)
# ---------------------------------------------------------------------------

def bc_symmetry():
    return dict(type=A.BC_SYMMETRY, kind=A.KIND_NONE, specified=A.SPEC_NONE, thermal=A.THERMAL_ADIABATIC)


def bc_wall(slip=False, T=None, qdot=None):
    d = dict(type=A.BC_WALL, kind=A.KIND_SLIP if slip else A.KIND_NONE, specified=A.SPEC_NONE,
             thermal=A.THERMAL_ADIABATIC)
    if not slip:
        d.update(V_fixed=1, V=(0.0, 0.0, 0.0))
    if T is not None:
        d.update(thermal=A.THERMAL_FIXED_T, T=T, T_fixed=1, specified=A.SPEC_T)
    if qdot is not None:
        d.update(thermal=A.THERMAL_FIXED_Q, qdot=qdot, qdot_fixed=1)
    return d


def bc_inlet_state(p, T, V, R):
    """velocity inlet with p and T given: specified = BC_STATE (ns_set_bcs.cc:52-74)."""
    return dict(type=A.BC_INLET, kind=A.KIND_VELOCITY, specified=A.SPEC_STATE, thermal=A.THERMAL_FIXED_T,
                p=p, T=T, rho=p / (R * T), V=tuple(V), p_fixed=1, T_fixed=1, rho_fixed=1, V_fixed=1)


def bc_inlet_T(T, V):
    """velocity inlet with only T given: specified = BC_T (ns_set_bcs.cc:86-97)."""
    return dict(type=A.BC_INLET, kind=A.KIND_VELOCITY, specified=A.SPEC_T, thermal=A.THERMAL_FIXED_T,
                T=T, V=tuple(V), T_fixed=1, V_fixed=1)


def bc_stagnation(p_total, T_total):
    return dict(type=A.BC_INLET, kind=A.KIND_STAGNATION, specified=A.SPEC_NONE, thermal=A.THERMAL_ADIABATIC,
                p_total=p_total, T_total=T_total)


def bc_outlet_p(p):
    return dict(type=A.BC_OUTLET, kind=A.KIND_NONE, specified=A.SPEC_P, thermal=A.THERMAL_ADIABATIC,
                p=p, p_fixed=1)


def bc_outlet_extrapolate():
    return dict(type=A.BC_OUTLET, kind=A.KIND_NONE, specified=A.SPEC_NONE, thermal=A.THERMAL_ADIABATIC)


def make_options(bcs, gamma=1.4, Mw=28.97, Pref=0.0, Tref=0.0, mu=0.0, lam=0.0, sutherland=None, Pr=None,
                 order=A.ORDER_SECOND, jac_order=A.ORDER_FIRST, limiter=A.LIMITER_VK,
                 flux=A.FLUX_AUSM_PLUS_UP, limiter_threshold=0.0, Minf=1.0, wdiss=0.3, bl_height=0.0,
                 grad=(A.GRAD_CURVILINEAR, A.GRAD_CURVILINEAR, A.GRAD_LSQR), interp=A.INTERP_WTLI,
                 stencil=2, loads_bcs=(), moment_center=(0.0, 0.0, 0.0), turbulence=A.TURB_NONE, Pr_t=0.9,
                 rans_model=A.RANS_SST, k_low=1e-10, k_high=1e8, omega_low=1e-2, visc_ratio_limit=1e8):
    """Defaults are the reference's (src/read_inputs.cc:47-54,70-75,110-134)."""
    o = A.Options()
    o.gamma, o.Mw, o.Pref, o.Tref = gamma, Mw, Pref, Tref
    if sutherland is not None:
        o.visc_model = A.VISC_SUTHERLAND
        o.sut_mu_ref, o.sut_T_ref, o.sut_S = sutherland
    else:
        o.visc_model, o.mu = A.VISC_CONSTANT, mu
    if Pr is not None:
        o.cond_model, o.Pr = A.COND_PRANDTL, Pr
    else:
        o.cond_model, o.lambda_ = A.COND_CONSTANT, lam
    o.order, o.jac_order, o.limiter, o.convective_flux = order, jac_order, limiter, flux
    o.limiter_threshold, o.Minf, o.wdiss, o.bl_height = limiter_threshold, Minf, wdiss, bl_height
    o.grad_hex, o.grad_prism, o.grad_other = grad
    o.interp_method, o.stencil_size, o.skewness_tol = interp, stencil, 0.99
    for q in range(3):
        o.moment_center[q] = moment_center[q]
    o.n_bc = len(bcs)
    for b, d in enumerate(bcs):
        bc = o.bc[b]
        bc.type, bc.kind, bc.specified, bc.thermal = d["type"], d["kind"], d["specified"], d["thermal"]
        for key in ("p_fixed", "T_fixed", "rho_fixed", "V_fixed", "qdot_fixed"):
            setattr(bc, key, int(d.get(key, 0)))
        for key in ("p", "T", "rho", "p_total", "T_total", "mdot", "qdot"):
            setattr(bc, key, float(d.get(key, 0.0)))
        V = d.get("V", (0.0, 0.0, 0.0))
        for q in range(3):
            bc.V[q] = V[q]
        bc.include_in_loads = int(b in loads_bcs)
        bc.nut = float(d.get("nut", 0.0))
        bc.turb_intensity = float(d.get("turb_intensity", 1e-4))
        bc.eddy_visc_ratio = float(d.get("eddy_visc_ratio", 0.1))
    o.turbulence, o.Pr_t = turbulence, Pr_t
    o.rans_model, o.k_low, o.k_high, o.omega_low, o.visc_ratio_limit = rans_model, k_low, k_high, omega_low, visc_ratio_limit
    return o


def freestream_state(n, p, T, V):
    W = np.empty((n, 5))
    W[:, 0] = p
    W[:, 1:4] = np.asarray(V)[None, :]
    W[:, 4] = T
    return W


class Case:
    def __init__(self, name, raw, opt, W0, meta):
        self.name, self.raw, self.opt, self.W0, self.meta = name, raw, opt, W0, meta

    @property
    def n_cells(self):
        return len(self.raw["cell_ptr"]) - 1


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------

def wedge_case(ni=96, nj=48, seed=1, flux=A.FLUX_ROE):
    """Inviscid Euler, gamma 1.4, Mach 2 over a 15 degree ramp; Roe, second-order
    lsqr gradients, minmod limiter; freestream with seeded +-1% uniform noise on
    rho, u, p."""
    raw = meshgen.wedge(ni, nj)
    Mw, g = 28.97, 1.4
    R = gas_R(Mw)
    p, T = 101325.0, 300.0
    U = 2.0 * np.sqrt(g * R * T)
    bcs = [bc_inlet_state(p, T, (U, 0.0, 0.0), R), bc_outlet_extrapolate(), bc_wall(slip=True),
           bc_wall(slip=True), bc_symmetry(), bc_symmetry()]
    opt = make_options(bcs, gamma=g, Mw=Mw, flux=flux, limiter=A.LIMITER_MINMOD,
                       grad=(A.GRAD_LSQR, A.GRAD_LSQR, A.GRAD_LSQR), Minf=2.0, loads_bcs=(2,))
    n = len(raw["cell_ptr"]) - 1
    rng = np.random.default_rng(seed)
    rho = p / (R * T) * (1.0 + rng.uniform(-0.01, 0.01, n))
    u = U * (1.0 + rng.uniform(-0.01, 0.01, n))
    pp = p * (1.0 + rng.uniform(-0.01, 0.01, n))
    W = freestream_state(n, p, T, (U, 0.0, 0.0))
    W[:, 0], W[:, 1], W[:, 4] = pp, u, pp / (rho * R)
    return Case("wedge", raw, opt, W, dict(p=p, T=T, U=U, R=R, mach=2.0))


def pie_case(nr=6, nt=12, nz=3, seed=4, jitter=0.1, viscous=True):
    """Axial flow through a cylinder whose innermost ring of hexahedra touches
    the axis (hexahedra with two repeated nodes, which the set-up turns into
    prisms): Mach 0.3 velocity inlet at the bottom, pressure outlet at the
    top, slip outer wall; Roe, curvilinear/lsqr gradients, VK; the
    freestream with seeded 1% noise."""
    raw = meshgen.pie(nr, nt, nz, jitter=jitter, seed=seed)
    Mw, g = 28.97, 1.4
    R = gas_R(Mw)
    p, T = 101325.0, 300.0
    U = 0.3 * np.sqrt(g * R * T)
    mu = 1.0e-3 if viscous else 0.0
    bcs = [bc_wall(slip=True), bc_inlet_velocity(A.SPEC_P, (0.0, 0.0, U), p=p), bc_outlet(A.SPEC_P, p=p)]
    opt = make_options(bcs, gamma=g, Mw=Mw, mu=mu, Pr=0.72 if viscous else None, flux=A.FLUX_ROE, limiter=A.LIMITER_VK,
                       Minf=0.3, loads_bcs=(0,))
    n = len(raw["cell_ptr"]) - 1
    rng = np.random.default_rng(seed)
    W = freestream_state(n, p, T, (0.0, 0.0, U))
    W[:, 0] *= 1.0 + rng.uniform(-0.01, 0.01, n)
    W[:, 3] *= 1.0 + rng.uniform(-0.01, 0.01, n)
    W[:, 4] *= 1.0 + rng.uniform(-0.01, 0.01, n)
    return Case("pie", raw, opt, W, dict(p=p, T=T, U=U, R=R, mach=0.3, mu=mu))


def cylinder_case(nt=256, nr=128, seed=2, noise=0.0):
    """Laminar compressible NS past a unit cylinder, Mach 0.2, Re_D = 40, Pr 0.72,
    Sutherland viscosity; AUSM+up, lsqr gradients for every cell type."""
    raw = meshgen.cylinder_ogrid(nt, nr, seed=seed)
    Mw, g = 28.97, 1.4
    R = gas_R(Mw)
    p, T = 101325.0, 300.0
    U = 0.2 * np.sqrt(g * R * T)
    rho = p / (R * T)
    mu_inf = rho * U * 1.0 / 40.0
    bcs = [bc_wall(slip=False), bc_inlet_T(T, (U, 0.0, 0.0)), bc_outlet_p(p), bc_symmetry(), bc_symmetry()]
    opt = make_options(bcs, gamma=g, Mw=Mw, sutherland=(mu_inf, T, 110.4), Pr=0.72, flux=A.FLUX_AUSM_PLUS_UP,
                       grad=(A.GRAD_LSQR, A.GRAD_LSQR, A.GRAD_LSQR), Minf=0.2, loads_bcs=(0,))
    n = len(raw["cell_ptr"]) - 1
    W = freestream_state(n, p, T, (U, 0.0, 0.0))
    if noise:
        rng = np.random.default_rng(seed + 100)
        W[:, 0] *= 1.0 + rng.uniform(-noise, noise, n)
        W[:, 1] *= 1.0 + rng.uniform(-noise, noise, n)
        W[:, 4] *= 1.0 + rng.uniform(-noise, noise, n)
    return Case("cylinder", raw, opt, W, dict(p=p, T=T, U=U, R=R, mach=0.2, mu=mu_inf))


def smooth_perturbation(xyz, seed, amp=0.05, modes=8):
    """Sum of `modes` random-phase Fourier modes (wave numbers 1..3 per axis),
    normalised to max |.| = 1, one field per variable; returns (n, 5)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((len(xyz), 5))
    for v in range(5):
        f = np.zeros(len(xyz))
        for _ in range(modes):
            kv = rng.integers(1, 4, size=3)
            ph = rng.uniform(0.0, 2.0 * np.pi, size=3)
            f += np.prod(np.sin(2.0 * np.pi * kv[None, :] * xyz + ph[None, :]), axis=1)
        out[:, v] = amp * f / max(np.abs(f).max(), 1e-300)
    return out


def bc_inlet_velocity(spec, V, p=0.0, T=0.0, rho=0.0):
    """velocity inlet with p, T or rho given (ns_set_bcs.cc:52-97)"""
    d = dict(type=A.BC_INLET, kind=A.KIND_VELOCITY, specified=spec, thermal=A.THERMAL_ADIABATIC, V=tuple(V), V_fixed=1,
             p=p, T=T, rho=rho)
    d.update({A.SPEC_P: dict(p_fixed=1), A.SPEC_T: dict(T_fixed=1), A.SPEC_RHO: dict(rho_fixed=1),
              A.SPEC_STATE: dict(p_fixed=1, T_fixed=1, rho_fixed=1)}.get(spec, {}))
    return d


def bc_mdot(spec, mdot, p=0.0, T=0.0, rho=0.0):
    """mass-flow inlet (ns_apply_bcs.cc:90-135): mdot over the region's area"""
    return dict(type=A.BC_INLET, kind=A.KIND_MDOT, specified=spec, thermal=A.THERMAL_ADIABATIC, mdot=mdot, p=p, T=T,
                rho=rho)


def bc_outlet(spec, kind=A.KIND_NONE, p=0.0, T=0.0, rho=0.0):
    """outlet with p, T, rho or nothing given, and a reverse-flow kind
    (ns_apply_bcs.cc:157-215)"""
    d = dict(type=A.BC_OUTLET, kind=kind, specified=spec, thermal=A.THERMAL_ADIABATIC, p=p, T=T, rho=rho)
    d.update({A.SPEC_P: dict(p_fixed=1), A.SPEC_T: dict(T_fixed=1), A.SPEC_RHO: dict(rho_fixed=1)}.get(spec, {}))
    return d


# Boundary-condition variants of a small 3-D channel (inlet at x = 0,
# outlet at x = L, the wall under test at y = 0, a slip wall at y = H,
# symmetry in z): each names (inlet, outlet, lower wall, flux).
BC_VARIANTS = {
    "vel_p__out_t__wall_fixed_t": ("vel_p", "out_t", "fixed_t", A.FLUX_ROE),
    "vel_rho__out_rho__wall_fixed_q": ("vel_rho", "out_rho", "fixed_q", A.FLUX_AUSM_PLUS_UP),
    "vel_t__out_damp_reverse__wall_adiabatic": ("vel_t", "out_damp", "adiabatic", A.FLUX_VAN_LEER),
    "mdot_p__out_no_reverse__wall_fixed_t": ("mdot_p", "out_noreverse", "fixed_t", A.FLUX_ROE),
    "mdot_t__out_extrapolate__wall_fixed_q": ("mdot_t", "out_none", "fixed_q", A.FLUX_AUSM_PLUS_UP),
    "mdot_rho__out_force_supersonic__wall_slip": ("mdot_rho", "out_force", "slip", A.FLUX_ROE),
    "mdot_state__out_p__wall_adiabatic": ("mdot_state", "out_p", "adiabatic", A.FLUX_SD_SLAU),
    "stagnation__out_t_no_reverse__wall_fixed_t": ("stag", "out_t_noreverse", "fixed_t", A.FLUX_SW),
}


def bc_variant_case(variant, ni=10, nj=6, nk=4, seed=12, sa=False, limiter=A.LIMITER_VK):
    """A jittered 1 x 0.5 x 0.25 hex channel, Mach ~0.2 viscous flow with a
    sheared inflow profile u = U (0.3 + cos(2 pi y / H)) (so some outlet and
    inlet faces see reverse flow) plus seeded 1% noise, under one of
    BC_VARIANTS; sa=True adds nu~."""
    inlet_k, outlet_k, wall_k, flux = BC_VARIANTS[variant]
    Lx, H, Wz = 1.0, 0.5, 0.25
    rng = np.random.default_rng(seed)
    X = np.zeros((ni + 1, nj + 1, nk + 1, 3))
    X[..., 0] = np.linspace(0, Lx, ni + 1)[:, None, None]
    X[..., 1] = np.linspace(0, H, nj + 1)[None, :, None]
    X[..., 2] = np.linspace(0, Wz, nk + 1)[None, None, :]
    h = np.array([Lx / ni, H / nj, Wz / nk])
    jit = rng.uniform(-0.15, 0.15, X.shape) * h
    jit[0, :, :, 0] = jit[-1, :, :, 0] = 0.0
    jit[:, 0, :, 1] = jit[:, -1, :, 1] = 0.0
    jit[:, :, 0, 2] = jit[:, :, -1, 2] = 0.0
    X = X + jit
    regions = [("inlet", lambda i, j, k: i == 0), ("outlet", lambda i, j, k: i == ni),
               ("wall", lambda i, j, k: j == 0), ("top", lambda i, j, k: j == nj),
               ("zmin", lambda i, j, k: k == 0), ("zmax", lambda i, j, k: k == nk)]
    raw = meshgen.structured_hex(X, regions, dimension=3)
    Mw, g = 28.97, 1.4
    R = gas_R(Mw)
    p, T = 1.0e5, 300.0
    U = 0.2 * np.sqrt(g * R * T)
    rho = p / (R * T)
    mdot = rho * U * H * Wz
    mu = 2.0e-3
    inlets = {
        "vel_p": bc_inlet_velocity(A.SPEC_P, (U, 0.0, 0.0), p=1.01 * p),
        "vel_t": bc_inlet_velocity(A.SPEC_T, (U, 0.0, 0.0), T=1.02 * T),
        "vel_rho": bc_inlet_velocity(A.SPEC_RHO, (U, 0.0, 0.0), rho=1.01 * rho),
        "mdot_p": bc_mdot(A.SPEC_P, mdot, p=1.01 * p),
        "mdot_t": bc_mdot(A.SPEC_T, mdot, T=1.02 * T),
        "mdot_rho": bc_mdot(A.SPEC_RHO, mdot, rho=1.01 * rho),
        "mdot_state": bc_mdot(A.SPEC_STATE, mdot, p=1.01 * p, T=1.02 * T, rho=1.01 * p / (R * 1.02 * T)),
        "stag": bc_stagnation(p * (1 + 0.5 * (g - 1) * 0.04) ** (g / (g - 1)), T * (1 + 0.5 * (g - 1) * 0.04)),
    }
    outlets = {
        "out_p": bc_outlet(A.SPEC_P, p=0.99 * p),
        "out_t": bc_outlet(A.SPEC_T, T=0.98 * T),
        "out_rho": bc_outlet(A.SPEC_RHO, rho=0.99 * rho),
        "out_none": bc_outlet(A.SPEC_NONE),
        "out_damp": bc_outlet(A.SPEC_P, A.KIND_DAMP_REVERSE, p=0.99 * p),
        "out_noreverse": bc_outlet(A.SPEC_P, A.KIND_NO_REVERSE, p=0.99 * p),
        "out_force": bc_outlet(A.SPEC_P, A.KIND_FORCE_SUPERSONIC, p=0.99 * p),
        "out_t_noreverse": bc_outlet(A.SPEC_T, A.KIND_NO_REVERSE, T=0.98 * T),
    }
    walls = {"adiabatic": bc_wall(), "fixed_t": bc_wall(T=1.1 * T), "fixed_q": bc_wall(qdot=5.0e3),
             "slip": bc_wall(slip=True)}
    inlet = dict(inlets[inlet_k])
    nut_inf = 3.0 * mu / rho
    inlet["nut"] = nut_inf
    bcs = [inlet, outlets[outlet_k], walls[wall_k], bc_wall(slip=True), bc_symmetry(), bc_symmetry()]
    opt = make_options(bcs, gamma=g, Mw=Mw, mu=mu, Pr=0.72, flux=flux, limiter=limiter, Minf=0.2, loads_bcs=(2,),
                       turbulence=A.TURB_SA if sa else A.TURB_NONE)
    n = len(raw["cell_ptr"]) - 1
    xc = meshgen.cell_centroids(raw)
    W = freestream_state(n, p, T, (U, 0.0, 0.0))
    W[:, 1] = U * (0.3 + np.cos(2 * np.pi * xc[:, 1] / H))
    W[:, 2] = 0.1 * U * np.sin(np.pi * xc[:, 0] / Lx)
    W[:, 0] *= 1.0 + rng.uniform(-0.01, 0.01, n)
    W[:, 1] *= 1.0 + rng.uniform(-0.01, 0.01, n)
    W[:, 4] *= 1.0 + rng.uniform(-0.01, 0.01, n)
    if sa:
        W = np.concatenate([W, (nut_inf * (1.0 + rng.uniform(-0.01, 0.01, n)))[:, None]], axis=1)
    return Case("bc_" + variant + ("_sa" if sa else ""), raw, opt, W, dict(p=p, T=T, U=U, R=R, mach=0.2, mu=mu, nut=nut_inf))


def smooth_state(cs, seed=21, amp=0.02):
    """Replace cs.W0 by the case's mean state times a seeded smooth
    perturbation (8 Fourier modes per variable over the bounding box), so that
    limiter switches (local extrema) are rare.  Returns cs."""
    xc = meshgen.cell_centroids(cs.raw)
    lo, hi = xc.min(axis=0), xc.max(axis=0)
    d = smooth_perturbation((xc - lo) / np.maximum(hi - lo, 1e-30), seed, amp=amp)
    W = cs.W0.mean(axis=0)[None, :] * np.ones((cs.n_cells, 1))
    W[:, 0] *= 1.0 + d[:, 0]
    W[:, 1] *= 1.0 + d[:, 1]
    W[:, 2] += cs.meta["U"] * d[:, 2]
    W[:, 3] += cs.meta["U"] * d[:, 3]
    W[:, 4] *= 1.0 + d[:, 4]
    if W.shape[1] == 6:   # SA nu~
        W[:, 5] *= 1.0 + smooth_perturbation((xc - lo) / np.maximum(hi - lo, 1e-30), seed + 7, amp=amp)[:, 0]
    cs.W0 = W
    return cs


def tetbox_case(n=64, seed=3, viscous=True):
    """Differentiation check: [0,1]^3 of n^3 hexes x 6 tets, Mach 0.5 freestream
    plus a seeded smooth 5% perturbation in rho, u, v, w, p; Roe, limited lsqr,
    viscous terms on."""
    raw = meshgen.tet_box(n)
    Mw, g = 28.97, 1.4
    R = gas_R(Mw)
    p, T = 101325.0, 300.0
    a = np.sqrt(g * R * T)
    U = 0.5 * a
    rho = p / (R * T)
    mu = rho * U * 1.0 / 1000.0 if viscous else 0.0
    bcs = [bc_inlet_state(p, T, (U, 0.0, 0.0), R), bc_outlet_p(p), bc_symmetry(), bc_symmetry(),
           bc_symmetry(), bc_symmetry()]
    opt = make_options(bcs, gamma=g, Mw=Mw, mu=mu, Pr=0.72 if viscous else None, flux=A.FLUX_ROE,
                       limiter=A.LIMITER_VK, limiter_threshold=5.0,
                       grad=(A.GRAD_LSQR, A.GRAD_LSQR, A.GRAD_LSQR), Minf=0.5)
    nc = len(raw["cell_ptr"]) - 1
    xc = meshgen.cell_centroids(raw)
    d = smooth_perturbation(xc, seed)
    rho_c = rho * (1.0 + d[:, 0])
    W = np.empty((nc, 5))
    W[:, 1] = U * (1.0 + d[:, 1])
    W[:, 2] = U * d[:, 2]
    W[:, 3] = U * d[:, 3]
    W[:, 0] = p * (1.0 + d[:, 4])
    W[:, 4] = W[:, 0] / (rho_c * R)
    return Case("tetbox", raw, opt, W, dict(p=p, T=T, U=U, R=R, mach=0.5, mu=mu))


def bump_case(ni=256, nj=96, nk=64, seed=4, noise=0.0, inlet="velocity"):
    """Transonic inviscid channel flow, Mach 0.85, 10% circular-arc bump.
    inlet="velocity": velocity inlet (V and T of the Mach 0.85 stream), the
    form used for the residual and tangent tests; the 10% bump chokes this
    channel (throat/inlet area 0.9 < A*/A = 0.991 at Mach 0.85), so with V
    fixed at the inlet there is no steady state and the inlet turns
    supersonic in pseudo-time.  inlet="stagnation": p0, T0 of the Mach 0.85
    stream, the well-posed choked-bump problem (the inlet Mach settles where
    the area ratio puts it) used for implicit solves to convergence."""
    raw = meshgen.bump_channel(ni, nj, nk)
    Mw, g = 28.97, 1.4
    R = gas_R(Mw)
    p, T = 101325.0, 300.0
    U = 0.85 * np.sqrt(g * R * T)
    if inlet == "stagnation":
        f = 1.0 + 0.5 * (g - 1.0) * 0.85 ** 2
        inl = bc_stagnation(p * f ** (g / (g - 1.0)), T * f)
    else:
        inl = bc_inlet_T(T, (U, 0.0, 0.0))
    bcs = [inl, bc_outlet_p(p), bc_wall(slip=True), bc_wall(slip=True),
           bc_symmetry(), bc_symmetry()]
    opt = make_options(bcs, gamma=g, Mw=Mw, flux=A.FLUX_ROE, limiter=A.LIMITER_VK, Minf=0.85, loads_bcs=(2,))
    n = len(raw["cell_ptr"]) - 1
    W = freestream_state(n, p, T, (U, 0.0, 0.0))
    if noise:
        rng = np.random.default_rng(seed)
        W[:, 0] *= 1.0 + rng.uniform(-noise, noise, n)
        W[:, 1] *= 1.0 + rng.uniform(-noise, noise, n)
        W[:, 4] *= 1.0 + rng.uniform(-noise, noise, n)
    return Case("bump", raw, opt, W, dict(p=p, T=T, U=U, R=R, mach=0.85))


def plate_case(ni=512, nj=128, nk=64, seed=5, noise=0.0, first=5.5e-6, sa=False, re=5.0e6, prism_frac=0.2):
    """Turbulent flat plate (Mach 0.3, Re_L = 5e6, or `re`), hex layers with
    the top 20% split into prisms, viscous Roe; sa=True adds the
    Spalart-Allmaras variable nu~ (freestream and inlet nu~ = 3 nu) as the
    sixth equation."""
    raw = meshgen.flat_plate(ni, nj, nk, first=first, prism_frac=prism_frac)
    Mw, g = 28.97, 1.4
    R = gas_R(Mw)
    p, T = 101325.0, 300.0
    U = 0.3 * np.sqrt(g * R * T)
    rho = p / (R * T)
    mu = rho * U * 1.0 / re
    nut_inf = 3.0 * mu / rho
    inlet = bc_inlet_T(T, (U, 0.0, 0.0))
    inlet["nut"] = nut_inf
    bcs = [inlet, bc_outlet_p(p), bc_wall(slip=False), bc_outlet_p(p), bc_symmetry(), bc_symmetry()]
    opt = make_options(bcs, gamma=g, Mw=Mw, mu=mu, Pr=0.72, flux=A.FLUX_ROE, limiter=A.LIMITER_VK,
                       Minf=0.3, loads_bcs=(2,), turbulence=A.TURB_SA if sa else A.TURB_NONE)
    n = len(raw["cell_ptr"]) - 1
    W = freestream_state(n, p, T, (U, 0.0, 0.0))
    if sa:
        W = np.concatenate([W, np.full((n, 1), nut_inf)], axis=1)
    if noise:
        rng = np.random.default_rng(seed)
        W[:, 0] *= 1.0 + rng.uniform(-noise, noise, n)
        W[:, 1] *= 1.0 + rng.uniform(-noise, noise, n)
        W[:, 4] *= 1.0 + rng.uniform(-noise, noise, n)
        if sa:
            W[:, 5] *= 1.0 + rng.uniform(-noise, noise, n)
    return Case("plate_sa" if sa else "plate", raw, opt, W, dict(p=p, T=T, U=U, R=R, mach=0.3, mu=mu, nut=nut_inf))


# ---------------------------------------------------------------------------
# segregated two-equation RANS (k-omega family)
# ---------------------------------------------------------------------------

def _kw_inflow(cs, intensity, visc_ratio):
    """freestream k and omega of a case (the inlet's definition,
    rans_set_bcs.cc:58-61, at its velocity and temperature)"""
    m = cs.meta
    R = m["R"]
    rho = m["p"] / (R * m["T"])
    k = max(1e-10, 1.5 * (intensity * m["U"]) ** 2)
    return k, rho * k / (visc_ratio * m["mu"])


def kw_state(cs, seed=31, amp=0.3, intensity=0.05, visc_ratio=10.0):
    """k and omega per cell (N x 2): the freestream values times smooth
    seeded variations (amplitude `amp`), so gradients, the blending and the
    cross-diffusion term are all non-trivial"""
    k0, w0 = _kw_inflow(cs, intensity, visc_ratio)
    xc = meshgen.cell_centroids(cs.raw)
    lo, hi = xc.min(axis=0), xc.max(axis=0)
    s = (xc - lo) / np.maximum(hi - lo, 1e-30)
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 6)
    fk = 1.0 + amp * np.sin(2 * np.pi * s[:, 0] + ph[0]) * np.cos(np.pi * s[:, 1] + ph[1]) + 0.3 * amp * np.sin(2 * np.pi * s[:, 2] + ph[2])
    fw = 1.0 + amp * np.cos(2 * np.pi * s[:, 1] + ph[3]) * np.sin(np.pi * s[:, 0] + ph[4]) + 0.3 * amp * np.cos(2 * np.pi * s[:, 2] + ph[5])
    return np.stack([k0 * fk, w0 * fw], axis=1)


def kw_case(variant="vel_p__out_t__wall_fixed_t", model=A.RANS_SST, ni=10, nj=6, nk=4, seed=12, intensity=0.05,
            visc_ratio=10.0, limiter=A.LIMITER_VK):
    """bc_variant_case with the segregated two-equation model `model` (the
    variant's flux, BCs and NS state); inlet turbulence intensity and eddy
    viscosity ratio as given.  cs.K0 holds an N x 2 k, omega state."""
    cs = bc_variant_case(variant, ni, nj, nk, seed=seed, limiter=limiter)
    o = cs.opt
    o.turbulence, o.rans_model = A.TURB_KW, model
    for b in range(o.n_bc):
        o.bc[b].turb_intensity, o.bc[b].eddy_visc_ratio = intensity, visc_ratio
    cs.name = "kw_%d_%s" % (model, variant)
    cs.K0 = kw_state(cs, intensity=intensity, visc_ratio=visc_ratio)
    cs.meta.update(intensity=intensity, visc_ratio=visc_ratio)
    return cs


def plate_kw_case(ni=512, nj=128, nk=64, model=A.RANS_SST, noise=0.0, intensity=0.01, visc_ratio=10.0, seed=5,
                  prism_frac=0.2):
    """The turbulent flat plate (Mach 0.3, Re_L 5e6, viscous Roe, top 20% of
    the layers split into prisms) with the segregated two-equation model:
    5 NS equations plus k and omega; freestream k, omega from the inlet's
    intensity and eddy viscosity ratio in cs.K0."""
    cs = plate_case(ni, nj, nk, seed=seed, noise=noise, prism_frac=prism_frac)
    o = cs.opt
    o.turbulence, o.rans_model = A.TURB_KW, model
    for b in range(o.n_bc):
        o.bc[b].turb_intensity, o.bc[b].eddy_visc_ratio = intensity, visc_ratio
    k0, w0 = _kw_inflow(cs, intensity, visc_ratio)
    cs.name = "plate_kw"
    cs.K0 = np.tile([k0, w0], (cs.n_cells, 1))
    cs.meta.update(intensity=intensity, visc_ratio=visc_ratio)
    return cs
