"""Converged supersonic wedge flows against oblique-shock theory.

A steady inviscid flow over a wedge of angle theta at Mach M carries an
attached oblique shock whose angle beta solves the theta-beta-M relation
    tan(theta) = 2 cot(beta) (M^2 sin^2 beta - 1) / (M^2 (gamma + cos 2 beta) + 2)
(weak root), with the pressure ratio p2/p1 = 1 + 2 gamma/(gamma+1)
(M^2 sin^2 beta - 1) behind it.  Newton's method on the GPU drives each case to
a steady state; the wall pressure behind the shock and the shock's angle,
read where the mid-pressure is crossed along a horizontal line, are checked
against those values.

Two cases:
* a synthetic Mach 2, 10 degree ramp (freecfd_amd.meshgen.wedge);
* the reference's shipped example (examples/wedge: Mach 2.5 over a 5 degree
  wedge, wedge.in), read from its CGNS mesh with freecfd_amd.cgns and run
  with the deck's values: inlet rho = 1, V = 935.414347, gauge p = 0 over
  the reference pressure 1e5; extrapolated outlet; symmetry; slip wall;
  viscosity 0; second order with the Venkatakrishnan limiter at threshold 1.
  Its mesh comes from tests/golden/wedge_example_mesh.npz, the example's
  wedge.cgns as the CGNS reader returns it (data only; test_cgns.py checks
  the fixture against the file where the reference tree is present)."""
import os

import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

def _oblique(M, theta_deg, g=1.4):
    th = np.radians(theta_deg)
    f = lambda b: 2.0 / np.tan(b) * (M * M * np.sin(b) ** 2 - 1) / (M * M * (g + np.cos(2 * b)) + 2) - np.tan(th)
    lo, hi = np.arcsin(1.0 / M) + 1e-12, np.radians(64.0)     # Mach angle .. beyond the weak root for these M
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if f(mid) < 0 else (lo, mid)
    b = 0.5 * (lo + hi)
    return np.degrees(b), 1.0 + 2.0 * g / (g + 1.0) * (M * M * np.sin(b) ** 2 - 1.0)


def test_oblique_relation_known_values():
    """Tabulated values (NACA Report 1135 charts): M 2, 15 deg: beta 45.3,
    p2/p1 2.19; M 2, 10 deg: beta 39.3, p2/p1 1.71; M 2.5, 5 deg: beta 27.4,
    p2/p1 1.38."""
    b, r = _oblique(2.0, 15.0)
    assert abs(b - 45.34) < 0.05 and abs(r - 2.195) < 0.005
    b, r = _oblique(2.0, 10.0)
    assert abs(b - 39.31) < 0.05 and abs(r - 1.707) < 0.005
    b, r = _oblique(2.5, 5.0)
    assert abs(b - 27.42) < 0.05 and abs(r - 1.380) < 0.005


def _solve(ns, W0, steps, cfl0, target):
    """Pseudo-time Newton steps with the example deck's CFL ramp (initial 5,
    growth 1.1, time marching in wedge.in) and a 20% cap on the relative
    change of p and T per step."""
    ns.set_state(W0)
    opts = A.newton_opts(cfl=cfl0, restart=30, max_iters=60, rtol=1e-3, precond=A.PC_BLOCK_SGS, sweeps=2,
                         max_rel_update=0.2)
    r = ns.solve(opts, max_steps=steps, drop=1e-8, cfl_initial=cfl0, cfl_growth=1.1, cfl_target=target,
                 callback=lambda k, st, cfl: print("  step %d cfl %.1f res %.3e" % (k, cfl, st.res_norm), flush=True)
                 if k % 10 == 0 else None)
    print("  %s in %d steps, drop %.2e" % (r["status"], r["steps"], r["drop"]), flush=True)
    return r


def _shock_point(cen, p, y, band, pmid, x_from):
    """(x, y) where p first rises through pmid along the cells within band of
    y, ordered by x (linear between the two cells that bracket it)."""
    sel = np.nonzero((np.abs(cen[:, 1] - y) < band) & (cen[:, 0] > x_from))[0]
    sel = sel[np.argsort(cen[sel, 0])]
    k = np.nonzero(p[sel] >= pmid)[0][0]
    assert k > 0
    a, b = sel[k - 1], sel[k]
    t = (pmid - p[a]) / (p[b] - p[a])
    return cen[a, :2] + t * (cen[b, :2] - cen[a, :2])


def _check(ns, wall_bc, p1, Pref, M, theta, corner, x_range, y_line, band):
    beta, ratio = _oblique(M, theta)
    m = ns.metrics()
    N = ns.info.n_cells
    W = ns.state_host(include_ghosts=False)
    cen = m["cell_centroid"].reshape(-1, 3)[:N]
    p = W[:, 0] + Pref
    wf = np.nonzero(m["face_bc"] == wall_bc)[0]
    par = m["face_parent"][wf]
    near = par[(cen[par, 0] > x_range[0]) & (cen[par, 0] < x_range[1])]
    got = p[near] / p1
    print("  wall p2/p1 %.4f .. %.4f (theory %.4f)" % (got.min(), got.max(), ratio), flush=True)
    assert np.abs(got - ratio).max() < 0.02 * ratio
    xs, ys = _shock_point(cen, p, y_line, band, 0.5 * (1.0 + ratio) * p1, corner[0])
    b = np.degrees(np.arctan2(ys - corner[1], xs - corner[0]))
    print("  shock angle %.2f deg (theory %.2f)" % (b, beta), flush=True)
    assert abs(b - beta) < 1.0


@pytest.mark.gpu
def test_synthetic_ramp_mach2_10deg():
    """A 10 degree ramp at x = 0.2 in a unit-high channel at Mach 2: the shock
    (39.3 degrees) leaves through the outlet before reaching the top wall, so
    the flow behind it stays uniform.  Numerics as in the example deck."""
    from freecfd_amd import meshgen
    from freecfd_amd.ns import NavierStokes
    raw = meshgen.wedge(160, 96, length=1.0, height=1.0, ramp_x=0.2, angle_deg=10.0)
    Mw, g = 28.97, 1.4
    R = cases.gas_R(Mw)
    p, T = 101325.0, 300.0
    U = 2.0 * np.sqrt(g * R * T)
    bcs = [cases.bc_inlet_state(p, T, (U, 0.0, 0.0), R), cases.bc_outlet_extrapolate(), cases.bc_wall(slip=True),
           cases.bc_wall(slip=True), cases.bc_symmetry(), cases.bc_symmetry()]
    opt = cases.make_options(bcs, gamma=g, Mw=Mw, limiter=A.LIMITER_VK, limiter_threshold=1.0, Minf=1.0)
    n = len(raw["cell_ptr"]) - 1
    print("case ramp 160x96 Mach 2, 10 deg: %d cells" % n, flush=True)
    ns = NavierStokes(raw, opt)
    r = _solve(ns, cases.freestream_state(n, p, T, (U, 0.0, 0.0)), 300, 5.0, 120.0)
    assert r["drop"] < 1e-5
    _check(ns, 2, p, 0.0, 2.0, 10.0, (0.2, 0.0), (0.3, 0.9), 0.3, 0.006)
    ns.close()


WEDGE_MESH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wedge_example_mesh.npz")


def shipped_wedge_mesh():
    """the reference's examples/wedge mesh as the CGNS reader returns it,
    from the committed fixture (tests/golden/make_wedge_example_mesh.py)"""
    d = np.load(WEDGE_MESH, allow_pickle=False)
    raw = {k: d[k] for k in ("xyz", "cell_ptr", "cell_nodes", "bc_ptr", "bc_nodes")}
    raw["dimension"] = int(d["dimension"])
    raw["bc_names"] = [str(x) for x in d["bc_names"]]
    return raw


@pytest.mark.gpu
def test_shipped_wedge_example_mach25_5deg():
    from freecfd_amd.ns import NavierStokes
    raw = shipped_wedge_mesh()
    assert raw["bc_names"] == ["inlet", "outlet", "symmetry", "wall"]
    Mw, g, Pref = 28.97, 1.4, 1.0e5                 # wedge.in: molar mass, gamma, reference p
    R = cases.gas_R(Mw)
    rho, V = 1.0, (935.414347, 0.0, 0.0)
    T = Pref / (rho * R)
    inlet = dict(type=A.BC_INLET, kind=A.KIND_VELOCITY, specified=A.SPEC_STATE, thermal=A.THERMAL_FIXED_T,
                 p=0.0, T=T, rho=rho, V=V, p_fixed=1, T_fixed=1, rho_fixed=1, V_fixed=1)
    bcs = [inlet, cases.bc_outlet_extrapolate(), cases.bc_symmetry(), cases.bc_wall(slip=True)]
    opt = cases.make_options(bcs, gamma=g, Mw=Mw, Pref=Pref, limiter=A.LIMITER_VK, limiter_threshold=1.0, Minf=1.0)
    n = len(raw["cell_ptr"]) - 1
    W0 = cases.freestream_state(n, 0.0, T, V)      # IC_1
    print("case shipped wedge example: %d cells, Mach %.4f" % (n, V[0] / np.sqrt(g * Pref / rho)), flush=True)
    ns = NavierStokes(raw, opt)
    r = _solve(ns, W0, 300, 5.0, 120.0)
    assert r["drop"] < 1e-5
    _check(ns, 3, Pref, Pref, 2.5, 5.0, (1.0, 0.0), (3.0, 10.0), 2.0, 0.06)
    ns.close()
