"""Checks on the configurations' full-size meshes.

- Tet box (64^3 x 6 = 1,572,864 tets, Roe, limited lsqr, viscous) and bump
  channel (256x96x64 = 1,572,864 hexes, Roe, VK) against the CPU restatement:
  the residual parity bar of test_residual_gpu (ghost states 1e-14,
  gradients 1e-11 of their scale, residual 1e-12 of the per-cell summation
  scale) on seeded perturbed states.  The restatement needs ~15 s per mesh.
- The 5,046,272-cell SA-RANS plate (the benchmark workload) against the CPU
  restatement, whose exact bucketed wall search gives the reference's
  brute-force distances bit for bit in seconds: every set-up metric
  bit-exact (the device's wall distances included), the residual parity bar
  above, and the FD Jacobian on 4,000 sampled faces and 4,000 sampled cells
  (1e-6 per block: the SA source's pow); plus size-independent properties:
  two residual evaluations are bitwise equal; the exact tangent is linear
  (J(a u + b v) = a Ju + b Jv to rounding); the adjoint is its transpose
  (<v, J u> = <J^T v, u> to 1e-11 of sum |v Ju|).
- The dot-product test on the full tet box as well.
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

SCALES = np.array([1e5, 100.0, 100.0, 100.0, 300.0, 1e-4])   # p, u, v, w, T, nu~


def _parity(cs):
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    print("case", cs.name, "N", cs.n_cells, flush=True)
    o = Oracle(cs.raw, cs.opt)
    Ro, Rabs = o.residual(cs.W0)
    Wo, Go, phio = o.fields()
    o.close()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    ns.assemble_residual()
    ns.status()
    Rg = ns.residual_host()
    Wg = ns.state_host(include_ghosts=True)
    Gg, phig = ns.gradients_limiters_host()
    ns.close()
    gscale = np.abs(Wo).max(axis=0)
    assert np.all(np.abs(Wg - Wo) <= 1e-14 * gscale + 1e-300), "ghost state mismatch"
    gerr = (np.abs(Gg - Go) / (np.abs(Go).max(axis=0) + 1e-300)).max()
    perr = np.abs(phig - phio).max()
    err = (np.abs(Rg - Ro) / (Rabs + 1e-300)).max()
    print("  gradients %.3e  limiters %.3e  max |dR|/S = %.3e" % (gerr, perr, err), flush=True)
    assert gerr <= 1e-11
    assert perr <= 1e-9
    assert err <= 1e-12


def test_tetbox_full_size_matches_oracle():
    _parity(cases.tetbox_case(64))


def test_bump_full_size_matches_oracle():
    _parity(cases.bump_case(noise=0.01))


def _dot(ns, N, nv, seed):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(N, nv)) * SCALES[:nv]
    v = rng.normal(size=(N, nv))
    Ju = ns.to_host(ns.jvp(ns.to_device(u), mode=A.JAC_JVP_FULL), nv)
    JTv = ns.to_host(ns.vjp(ns.to_device(v)), nv)
    lhs, rhs = float(np.sum(v * Ju)), float(np.sum(JTv * u))
    rel = abs(lhs - rhs) / float(np.sum(np.abs(v * Ju)))
    print("  <v,Ju> = %.15e  <J^T v,u> = %.15e  rel %.2e" % (lhs, rhs, rel), flush=True)
    return rel


def test_tetbox_full_size_dot_product():
    from freecfd_amd.ns import NavierStokes
    cs = cases.tetbox_case(64)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    assert _dot(ns, cs.n_cells, 5, 31) < 1e-11
    ns.close()


def test_sa_plate_full_size_properties():
    from freecfd_amd.ns import NavierStokes
    cs = cases.smooth_state(cases.plate_case(sa=True))
    N = cs.n_cells
    print("case", cs.name, "N", N, flush=True)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    R1 = ns.assemble_residual().clone()
    R2 = ns.assemble_residual().clone()
    ns.status()
    assert bool((R1 == R2).all()), "residual not deterministic"
    # linearity of the exact tangent
    rng = np.random.default_rng(7)
    u = rng.normal(size=(N, 6)) * SCALES
    v = rng.normal(size=(N, 6)) * SCALES
    a, b = 0.75, -1.5
    Ju = ns.jvp(ns.to_device(u)).clone()
    Jv = ns.jvp(ns.to_device(v)).clone()
    Jw = ns.jvp(ns.to_device(a * u + b * v)).clone()
    lin = float((Jw - (a * Ju + b * Jv)).abs().max() / (a * Ju).abs().max())
    print("  tangent linearity: max |J(au+bv) - aJu - bJv| / max|aJu| = %.2e" % lin, flush=True)
    assert lin < 1e-10
    assert _dot(ns, N, 6, 11) < 1e-11
    ns.close()


def test_sa_plate_full_size_matches_oracle():
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = cases.plate_case(sa=True, noise=0.01)
    N = cs.n_cells
    print("case", cs.name, "N", N, flush=True)
    o = Oracle(cs.raw, cs.opt)
    mo = o.metrics()
    Ro, Rabs = o.residual(cs.W0, with_jacobian=True)
    Wo, Go, phio = o.fields()
    JLo, JRo, Do = o.jacobian()
    o.close()
    print("  oracle done", flush=True)
    ns = NavierStokes(cs.raw, cs.opt)
    mg = ns.metrics()
    for key, a in mo.items():
        if key != "cell_perm":
            assert a.shape == mg[key].shape and np.array_equal(a, mg[key]), key
    del mo, mg
    print("  set-up metrics bit-exact (wall distance from the device)", flush=True)
    ns.set_state(cs.W0)
    ns.assemble_residual()
    ns.status()
    Rg = ns.residual_host()
    Wg = ns.state_host(include_ghosts=True)
    Gg, phig = ns.gradients_limiters_host()
    gscale = np.abs(Wo).max(axis=0)
    assert np.all(np.abs(Wg - Wo) <= 1e-14 * gscale + 1e-300), "ghost state mismatch"
    gerr = (np.abs(Gg - Go) / (np.abs(Go).max(axis=0) + 1e-300)).max()
    perr = np.abs(phig - phio).max()
    err = (np.abs(Rg - Ro) / (Rabs + 1e-300)).max()
    print("  gradients %.3e  limiters %.3e  max |dR|/S = %.3e" % (gerr, perr, err), flush=True)
    assert gerr <= 1e-11 and perr <= 1e-9 and err <= 1e-12
    ns.fd_jacobian()
    ns.status()
    JLg, JRg, Dg = ns.jacobian_host()
    rng = np.random.default_rng(17)
    faces = rng.choice(len(JLo), 4000, replace=False)
    cells = rng.choice(N, 4000, replace=False)
    def blockwise(a, b):
        keep = np.abs(b).reshape(len(b), -1).max(axis=1) > 0
        a, b = a[keep], b[keep]
        return float((np.abs(a - b).reshape(len(b), -1).max(axis=1) / np.abs(b).reshape(len(b), -1).max(axis=1)).max())
    e = (blockwise(JLg[faces], JLo[faces]), blockwise(JRg[faces], JRo[faces]), blockwise(Dg[cells], Do[cells]))
    print("  FD Jacobian on 4000 faces / cells: JL %.2e  JR %.2e  D %.2e" % e, flush=True)
    assert max(e) <= 1e-6
    ns.close()
