"""Differentiation checks of the residual through a torch.autograd.Function
(tests/residual_autograd.py) over fcfd_jvp / fcfd_vjp.

Small mesh: the wrapper against torch.autograd.gradcheck (fast mode) on a
nondimensionalised state, forward-mode AD against central differences,
reverse mode against the dot-product identity.

Full size, the differentiation-check configuration: [0,1]^3 of 64^3 hexes x
6 tetrahedra = 1,572,864 cells, Mach 0.5 plus a seeded smooth 5%
perturbation, Roe, limited lsqr gradients, viscous terms on; tangent and
seed vectors N(0,1) per cell and variable (times the variable scales):
  - the exact tangent (forward-mode AD) against central differences of the
    residual, h = 1e-7: the cells whose limiter switches inside +-h (a
    kink of min/max) are few, and the median error is at the FD level;
  - <v, J u> (forward mode) against <J^T v, u> (reverse mode);
  - the frozen-gradient tangent against the reference-style FD Jacobian of
    the CPU restatement, column by column on 1,000 seeded random cells
    (5,000 columns; cells whose face neighbourhoods do not overlap, so the
    columns of one variable come from one product).
"""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

SCALES = np.array([1e5, 100.0, 100.0, 100.0, 300.0, 1e-4])


def _ns(cs):
    from freecfd_amd.ns import NavierStokes
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    return ns


def _interior(ns, X_host):
    """N x nv reference-order host array -> [nv][N] device tensor (internal order)"""
    return ns.to_device(X_host)[:, :ns.info.n_cells].clone()


def _host(ns, X):
    """[nv][N] device tensor -> N x nv reference-order host array"""
    import residual_autograd as RA
    return ns.to_host(RA._pad(ns, X), ns.nv)


def test_autograd_wrapper_small_mesh():
    import torch
    import residual_autograd as RA
    cs = cases.smooth_state(cases.tetbox_case(3))
    cs.opt.limiter = A.LIMITER_NONE
    print("case", cs.name, "N", cs.n_cells, flush=True)
    ns = _ns(cs)
    W = _interior(ns, cs.W0)
    S = torch.tensor(SCALES[:5], dtype=torch.float64, device=W.device)[:, None]
    R0 = RA.residual(W, ns)
    Rs = float(R0.abs().max())
    # gradcheck on Z = W / S, outputs R / max|R|: forward and reverse mode
    Z = (W / S).detach().requires_grad_(True)
    f = lambda z: RA.residual(z * S, ns) / Rs
    assert torch.autograd.gradcheck(f, (Z,), eps=1e-7, atol=1e-6, rtol=1e-4, fast_mode=True,
                                    check_forward_ad=True, check_backward_ad=True, check_undefined_grad=False,
                                    check_batched_grad=False)
    rng = np.random.default_rng(2)
    u = torch.tensor(rng.normal(size=W.shape), device=W.device) * S
    v = torch.tensor(rng.normal(size=W.shape), device=W.device)
    _, Ju = RA.jvp(ns, W, u)
    JTv = RA.vjp(ns, W, v)
    lhs, rhs = float((v * Ju).sum()), float((JTv * u).sum())
    rel = abs(lhs - rhs) / float((v * Ju).abs().sum())
    print("  dot test through autograd: %.2e" % rel, flush=True)
    assert rel < 1e-11


def _select_cells(m, N, count, seed):
    """`count` random interior cells whose closed face neighbourhoods
    (the cell and the cells across its faces) are pairwise disjoint"""
    par, nei = m["face_parent"], m["face_neighbor"]
    nb = [[] for _ in range(N)]
    for p, n in zip(par, nei):
        if n < N:
            nb[p].append(n)
            nb[n].append(p)
    rng = np.random.default_rng(seed)
    taken = np.zeros(N, bool)
    out = []
    for c in rng.permutation(N):
        hood = [c] + nb[c]
        if taken[hood].any():
            continue
        taken[hood] = True
        out.append(c)
        if len(out) == count:
            break
    return np.array(out), nb


def test_config3_full_size_tangent_adjoint_and_fd_columns():
    import torch
    import residual_autograd as RA
    from oracle.oracle import Oracle
    cs = cases.tetbox_case(64)
    N = cs.n_cells
    print("case", cs.name, "N", N, flush=True)
    ns = _ns(cs)
    W = _interior(ns, cs.W0)
    rng = np.random.default_rng(4)
    u_h = rng.normal(size=(N, 5)) * SCALES[:5]
    v_h = rng.normal(size=(N, 5))
    u, v = _interior(ns, u_h), _interior(ns, v_h)
    # 1. exact tangent vs central differences, h = 1e-7: first without the
    #    limiter (smooth everywhere), then as configured (limited lsqr)
    def fd_errors(ns_, W_):
        _, Ju_ = RA.jvp(ns_, W_, u)
        h = 1e-7
        fd = (RA.residual(W_ + h * u, ns_) - RA.residual(W_ - h * u, ns_)) / (2 * h)
        scale = fd.abs().amax(dim=1, keepdim=True)
        return ((Ju_ - fd).abs() / scale).amax(dim=0).cpu().numpy(), Ju_
    cs_u = cases.tetbox_case(64)
    cs_u.opt.limiter = A.LIMITER_NONE
    ns_u = _ns(cs_u)
    per_cell, _ = fd_errors(ns_u, _interior(ns_u, cs_u.W0))
    ns_u.close()
    print("  unlimited: JVP vs central FD per cell: median %.2e, max %.2e" % (np.median(per_cell), per_cell.max()), flush=True)
    # (a few faces sit where Roe switches its one-sided form, u_n = 0: a kink)
    assert np.median(per_cell) < 1e-8 and per_cell.max() < 1e-4 and np.mean(per_cell > 1e-6) < 1e-4
    per_cell, Ju = fd_errors(ns, W)
    bad = float(np.mean(per_cell > 1e-5))
    print("  limited: median %.2e, fraction of cells above 1e-5 (a limiter switch inside +-h) %.2e"
          % (np.median(per_cell), bad), flush=True)
    assert np.median(per_cell) < 1e-6 and bad < 0.02
    # 2. dot-product test, forward vs reverse mode
    JTv = RA.vjp(ns, W, v)
    lhs, rhs = float((v * Ju).sum()), float((JTv * u).sum())
    rel = abs(lhs - rhs) / float((v * Ju).abs().sum())
    print("  <v,Ju> %.15e  <J^T v,u> %.15e  rel %.2e" % (lhs, rhs, rel), flush=True)
    assert rel < 1e-11
    # 3. frozen tangent vs the restatement's FD Jacobian on 1,000 cells
    o = Oracle(cs.raw, cs.opt)
    o.residual(cs.W0, with_jacobian=True)
    JL, JR, D = o.jacobian()
    m = o.metrics()
    o.close()
    cells, nb = _select_cells(m, N, 1000, seed=5)
    assert len(cells) == 1000
    faces_of = [[] for _ in range(N)]
    for f, (p, n) in enumerate(zip(m["face_parent"], m["face_neighbor"])):
        faces_of[p].append(f)
        if n < N:
            faces_of[n].append(f)
    worst = 0.0
    for var in range(5):
        e = np.zeros((N, 5))
        e[cells, var] = 1.0
        col = ns.to_host(ns.jvp(ns.to_device(e), mode=A.JAC_JVP_FROZEN), 5)
        for j in cells:
            ref = {j: np.zeros(5)}
            for f in faces_of[j]:
                p, n = m["face_parent"][f], m["face_neighbor"][f]
                blk = JL[f] if p == j else JR[f]          # [var][eq]
                ref[p] = ref.get(p, np.zeros(5)) + blk[var]
                if n < N:
                    ref[n] = ref.get(n, np.zeros(5)) - blk[var]
            rows = sorted(ref)
            R_ref = np.array([ref[r] for r in rows])
            R_got = col[rows]
            err = np.abs(R_got - R_ref).max() / np.abs(R_ref).max()
            worst = max(worst, err)
    print("  frozen JVP vs FD Jacobian, 5000 columns: worst column %.2e" % worst, flush=True)
    # the reference FD step (1.5e-7 absolute) is ~1e-12 of p ~ 1e5, so the FD
    # columns carry ~1e-4 relative noise of their own
    assert worst < 1e-3
    ns.close()
