"""Pins of the Spalart-Allmaras restatement in the CPU oracle (build-only
model: SA without ft2, with the S~ modification and the negative-nu~ branch
of Allmaras, Johnson & Spalart, ICCFD7-1902, 2012).

  - fv1(chi = cv1) = 1/2 (mu_t = rho nu~ / 2 there);
  - the log layer: nu~ = kappa u_tau d with dU/dy = u_tau/(kappa d) is an exact
    solution of the high-Re model (the calibration cw1 = cb1/kappa^2 +
    (1 + cb2)/sigma): source/rho + (1/sigma)(kappa u_tau)^2 = 0 up to the
    laminar terms (O(nu/nu~));
  - nu~ = 0: no source, no eddy viscosity; continuity of the source across
    nu~ = 0 (standard and negative branches) and across the S~ switch;
  - residual level: uniform state in a closed slip box gives R = 0 in all six
    equations; with V = 0 the nu~ fluxes of a Cartesian closed slip box
    telescope (sum of R6 - V S over cells = 0);
  - FD Jacobian (SA rows and the source block) vs central differences of the
    whole residual with gradients, limiters and ghosts frozen, on interior cells.
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases, meshgen

KAPPA, SIGMA, CB1, CB2, CV1, CV2 = 0.41, 2.0 / 3.0, 0.1355, 0.622, 7.1, 0.7


def _probe(rho, mu, nut, dudy=0.0, dnut=(0.0, 0.0, 0.0), d=1.0):
    from oracle.oracle import sa_probe
    gvel = np.zeros((3, 3))
    gvel[0, 1] = dudy
    return sa_probe(rho, mu, nut, gvel, np.asarray(dnut), d)


def test_fv1_half_at_cv1(oracle_lib):
    rho, mu = 1.2, 1.8e-5
    nut = CV1 * mu / rho
    r = _probe(rho, mu, nut)
    assert abs(r["mu_t"] - 0.5 * rho * nut) <= 1e-15 * rho * nut


@pytest.mark.parametrize("d", [1e-4, 1e-3, 1e-2])
def test_log_layer_is_exact_solution(d, oracle_lib):
    rho, utau = 1.2, 0.5
    nut = KAPPA * utau * d
    mu = rho * nut * 1e-9            # chi = 1e9: fv2 -> 0, fv1 -> 1
    r = _probe(rho, mu, nut, dudy=utau / (KAPPA * d), dnut=(0.0, KAPPA * utau, 0.0), d=d)
    diffusion = (KAPPA * utau) ** 2 / SIGMA        # (1/sigma) d/dy(nu~ dnu~/dy) for the linear profile
    bal = r["source"] / rho + diffusion
    assert abs(bal) <= 1e-7 * diffusion


def test_zero_nut(oracle_lib):
    r = _probe(1.2, 1.8e-5, 0.0, dudy=100.0, d=1e-3)
    assert r["source"] == 0.0 and r["mu_t"] == 0.0 and r["fn"] == 1.0


def test_continuity_across_zero_and_negative_branch(oracle_lib):
    rho, mu, d = 1.2, 1.8e-5, 1e-3
    a = _probe(rho, mu, 1e-14, dudy=100.0, d=d)["source"]
    b = _probe(rho, mu, -1e-14, dudy=100.0, d=d)["source"]
    ref = abs(_probe(rho, mu, 1e-6, dudy=100.0, d=d)["source"])
    assert abs(a) < 1e-6 * ref and abs(b) < 1e-6 * ref
    neg = _probe(rho, mu, -1e-6, dudy=100.0, d=d)
    assert neg["mu_t"] == 0.0
    # negative branch: P = cb1 (1 - ct3) Omega nu~ > 0 (ct3 > 1, nu~ < 0) and
    # D = -cw1 (nu~/d)^2 < 0, so the source drives nu~ back up towards zero
    assert neg["P"] > 0.0 and neg["D"] < 0.0 and neg["source"] > 0.0
    assert abs(_probe(rho, mu, -1e-12, d=d)["fn"] - 1.0) < 1e-12


def test_s_tilde_switch_is_continuous(oracle_lib):
    """S~ = Omega + Sbar for Sbar >= -cv2 Omega and the Allmaras rational form
    below; both give 0.3 Omega at the switch, so P = cb1 S~ nu~ is continuous
    in d across it (fv2 < 0 for chi in (1, ~6))."""
    rho, mu = 1.0, 1.0
    nut, Om = 3.0, 1.0
    chi = nut
    fv1 = chi ** 3 / (chi ** 3 + CV1 ** 3)
    fv2 = 1.0 - chi / (1.0 + chi * fv1)
    assert fv2 < 0
    d_sw = np.sqrt(nut * fv2 / (-CV2 * Om) / KAPPA ** 2)   # Sbar = -cv2 Omega
    Ps = [_probe(rho, mu, nut, dudy=Om, d=d_sw * (1 + s))["P"] for s in (-1e-9, 1e-9)]
    assert abs(Ps[0] - Ps[1]) <= 1e-7 * abs(Ps[0])
    assert abs(Ps[0] / (CB1 * nut) - 0.3 * Om) <= 1e-6 * Om


def _sa_box(nx=5, ny=4, nz=3, slip=True):
    X = np.zeros((nx + 1, ny + 1, nz + 1, 3))
    X[..., 0] = np.linspace(0, 1.0, nx + 1)[:, None, None]
    X[..., 1] = np.linspace(0, 0.8, ny + 1)[None, :, None]
    X[..., 2] = np.linspace(0, 0.6, nz + 1)[None, None, :]
    bcs = [("xmin", lambda i, j, k: i == 0), ("xmax", lambda i, j, k: i == nx),
           ("ymin", lambda i, j, k: j == 0), ("ymax", lambda i, j, k: j == ny),
           ("zmin", lambda i, j, k: k == 0), ("zmax", lambda i, j, k: k == nz)]
    raw = meshgen.structured_hex(X, bcs, dimension=3)
    opt = cases.make_options([cases.bc_wall(slip=True)] * 6, flux=A.FLUX_ROE, mu=1.8e-5, Pr=0.72,
                             turbulence=A.TURB_SA, grad=(A.GRAD_LSQR,) * 3)
    return raw, opt


def test_uniform_state_closed_box_zero_residual(oracle_lib):
    from oracle.oracle import Oracle
    raw, opt = _sa_box()
    o = Oracle(raw, opt)
    assert o.info.nv == 6
    W = np.concatenate([cases.freestream_state(o.info.n_cells, 1e5, 300.0, (0.0, 0.0, 0.0)),
                        np.full((o.info.n_cells, 1), 4e-5)], axis=1)
    R, Rabs = o.residual(W)
    assert np.abs(R[:, :5]).max() <= 1e-12 * Rabs[:, :5].max()
    # nu~ row: only the source remains, no walls -> d = 1e20 -> S~ = 0, P = 0, D ~ 0
    assert np.abs(R[:, 5]).max() <= 1e-20


def test_nut_fluxes_telescope_at_rest(oracle_lib):
    """V = 0 (no mass flux), arbitrary nu~ field, Cartesian closed slip box:
    the diffusive nu~ fluxes cancel pairwise and vanish on the walls."""
    from oracle.oracle import Oracle
    raw, opt = _sa_box()
    opt.order = A.ORDER_FIRST
    o = Oracle(raw, opt)
    N = o.info.n_cells
    rng = np.random.default_rng(0)
    W = np.concatenate([cases.freestream_state(N, 1e5, 300.0, (0.0, 0.0, 0.0)),
                        (1.0 + rng.uniform(0, 1, (N, 1))) * 1e-4], axis=1)
    R, Rabs = o.residual(W)
    # subtract the sources: at rest with no walls they vanish (Omega = 0, d = 1e20) except the cb2 term
    _, G, _ = o.fields()
    xv = o.metrics()["cell_volume"][:N]
    rho = 1e5 / (cases.gas_R(28.97) * 300.0)
    src = xv * CB2 / SIGMA * rho * (G[:N, 15:18] ** 2).sum(axis=1)
    tot = (R[:, 5] - src).sum()
    assert abs(tot) <= 1e-12 * Rabs[:, 5].sum()


def test_sa_fd_jacobian_vs_frozen_central_differences(oracle_lib):
    """Diagonal blocks of the FD Jacobian (reference epsilon policy, with the
    nu~ step scaled by the inlet nu~, and the SA source block) vs central
    differences of the whole residual with gradients, limiters and ghost
    states frozen, on cells away from the boundary (where the FD Jacobian's BC
    re-application does not enter).  The state carries nonzero v, w so that no
    |V.n| kink sits at the linearisation point.  The one-sided steps (1.5e-7
    relative in nu~) leave rounding noise of up to ~5e-3 of a row's scale in
    the momentum/energy entries of the nu~ column; the nu~ row (source and
    nu~ fluxes) is held to 1e-6."""
    from oracle.oracle import Oracle, lib, _p
    cs = cases.plate_case(8, 8, 3, first=1e-3, noise=0.01, sa=True)
    cs.opt.order = A.ORDER_FIRST
    U = cs.meta["U"]
    rng = np.random.default_rng(7)
    cs.W0[:, 2] = U * 0.05 * rng.uniform(-1, 1, cs.n_cells)
    cs.W0[:, 3] = U * 0.05 * rng.uniform(-1, 1, cs.n_cells)
    o = Oracle(cs.raw, cs.opt)
    N = o.info.n_cells
    o.residual(cs.W0, with_jacobian=True)
    _, _, Dg = o.jacobian()
    o.prepare()
    Wfull, _, _ = o.fields()
    m = o.metrics()
    interior = np.ones(N, bool)
    for f in range(o.info.n_faces):
        if m["face_bc"][f] >= 0:
            interior[m["face_parent"][f]] = False
    cells = np.nonzero(interior)[0][:8]
    assert len(cells) > 0
    scale = np.abs(cs.W0).mean(axis=0)
    for c in cells:
        J = np.zeros((6, 6))
        for v in range(6):
            h = 1e-6 * scale[v]
            Rs = []
            for sgn in (1, -1):
                Wp = Wfull.copy()
                Wp[c, v] += sgn * h
                o.set_full_state(Wp)
                lib().orc_assemble(o.h, 0)
                R = np.zeros((N, 6))
                lib().orc_get_residual(o.h, _p(R), None)
                Rs.append(R[c])
            J[:, v] = -(Rs[0] - Rs[1]) / (2 * h)          # Dg holds -dR/dW
        rel = np.abs(Dg[c] - J) / np.abs(J).max(axis=1, keepdims=True)
        assert rel[:5, :5].max() < 1e-4, (c, rel)
        assert rel[:5, 5].max() < 1e-2, (c, rel)     # mu_t coupling, 1.5e-7 relative steps in nu~
        assert rel[5].max() < 1e-6, (c, rel)
