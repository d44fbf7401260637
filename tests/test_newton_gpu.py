"""GPU implicit step: the GMRES solve of the FD-Jacobian system, with each
preconditioner (block Jacobi, block SGS, aggregation multigrid), against a
direct sparse solve of the same system assembled from the CPU restatement's
Jacobian blocks and time terms (the preconditioner changes the iterations,
not the solution); convergence of the wedge case (config 1 numerics: Roe,
lsqr, minmod, CFL 10); and the implicit solve of the laminar cylinder to a
1e-10 residual drop with the multigrid preconditioner, FD/FP64 and
exact-tangent/FP32."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu


def _assemble(o, dt):
    """Reference implicit operator impOP (MatSetValues pattern of
    ns_assemble_linear_system.cc:144-189 + ns_time_terms.cc:147-163)."""
    m = o.metrics()
    JL, JR, Dg = o.jacobian()
    N = o.info.n_cells
    TT = o.time_terms(dt)
    rows, cols, vals = [], [], []
    for c in range(N):
        blk = Dg[c] + TT[c]
        for i in range(5):
            for j in range(5):
                rows.append(5 * c + i); cols.append(5 * c + j); vals.append(blk[i, j])
    for f in range(o.info.n_faces):
        if m["face_bc"][f] >= 0:
            continue
        p, n = m["face_parent"][f], m["face_neighbor"][f]
        for var in range(5):
            for j in range(5):
                rows.append(5 * p + j); cols.append(5 * n + var); vals.append(-JR[f, var, j])
                rows.append(5 * n + j); cols.append(5 * p + var); vals.append(JL[f, var, j])
    return sp.csr_matrix((vals, (rows, cols)), shape=(5 * N, 5 * N))


@pytest.mark.parametrize("fp32,precond", [(False, A.PC_BLOCK_JACOBI), (True, A.PC_BLOCK_JACOBI), (False, A.PC_BLOCK_SGS),
                                           (False, A.PC_MULTIGRID), (True, A.PC_MULTIGRID)])
def test_newton_linear_solve_matches_direct(fp32, precond):
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = cases.wedge_case(24, 12)
    print("case wedge 24x12 fp32=%s precond=%d" % (fp32, precond), flush=True)
    o = Oracle(cs.raw, cs.opt)
    R, _ = o.residual(cs.W0, with_jacobian=True)
    dt, _ = o.time_step(A.DT_CFL_LOCAL, 10.0)
    Amat = _assemble(o, dt)
    dW_ref = spla.spsolve(Amat.tocsc(), R.reshape(-1)).reshape(-1, 5)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    opts = A.newton_opts(cfl=10.0, fp32=fp32, restart=60, max_iters=400, rtol=1e-12 if not fp32 else 1e-6, abstol=0.0,
                         precond=precond)
    st = ns.newton_step(opts)
    W1 = ns.state_host(include_ghosts=False)
    dW = W1 - cs.W0
    scale = np.abs(dW_ref).max() + 1e-300
    err = (np.abs(dW - dW_ref) / scale).max()
    print("  iters %d rnorm %.3e/%.3e  max rel err %.3e" % (st.krylov_iters, st.krylov_rnorm, st.krylov_r0norm, err), flush=True)
    assert err < (1e-8 if not fp32 else 1e-4)
    assert abs(st.res_norm - np.linalg.norm(R)) <= 1e-12 * np.linalg.norm(R) + 1e-300 or np.isclose(st.res_norm, np.linalg.norm(R), rtol=1e-10)


def test_multigrid_linear_solve_matches_direct_with_large_coarse_level():
    """The multigrid-preconditioned solve on the 128x64 cylinder (8,192 cells),
    whose first coarse level is above the 1024 cells that one workgroup
    smooths in one launch (so it runs one launch per colour, the smaller
    levels one launch per sweep), against the direct sparse solve of the
    CPU restatement's system.  The case is viscous with Sutherland's law,
    whose pow is not bitwise equal on the two sides: the FD Jacobians agree
    to about 1e-6, and so do the solutions."""
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(128, 64, noise=0.01)
    print("case cylinder 128x64 multigrid", flush=True)
    o = Oracle(cs.raw, cs.opt)
    R, _ = o.residual(cs.W0, with_jacobian=True)
    dt, _ = o.time_step(A.DT_CFL_LOCAL, 10.0)
    Amat = _assemble(o, dt)
    dW_ref = spla.spsolve(Amat.tocsc(), R.reshape(-1)).reshape(-1, 5)
    o.close()
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    opts = A.newton_opts(cfl=10.0, restart=60, max_iters=400, rtol=1e-12, abstol=0.0, precond=A.PC_MULTIGRID, sweeps=1)
    st = ns.newton_step(opts)
    dW = ns.state_host(include_ghosts=False) - cs.W0
    err = (np.abs(dW - dW_ref) / (np.abs(dW_ref).max() + 1e-300)).max()
    print("  iters %d rnorm %.3e/%.3e  max rel err %.3e" % (st.krylov_iters, st.krylov_rnorm, st.krylov_r0norm, err), flush=True)
    assert st.krylov_converged
    assert err < 1e-5
    ns.close()


def test_multigrid_level_launch_sweeps_match_per_colour_launches():
    """The one-launch sweeps of the small coarse levels (all colours in one
    workgroup, barriers between colours) give bit-identical Newton steps to
    one launch per colour (FCFD_MG_NO_LEVEL_LAUNCH=1), over three steps with
    both Krylov precisions."""
    import os
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(64, 32)
    out = {}
    for mode in ("per_colour", "level"):
        if mode == "per_colour":
            os.environ["FCFD_MG_NO_LEVEL_LAUNCH"] = "1"
        else:
            os.environ.pop("FCFD_MG_NO_LEVEL_LAUNCH", None)
        ns = NavierStokes(cs.raw, cs.opt)
        ns.set_state(cs.W0)
        for fp32, jac in ((False, A.JAC_FD), (True, A.JAC_JVP_FULL), (False, A.JAC_FD)):
            opts = A.newton_opts(cfl=100.0, jacobian=jac, fp32=fp32, restart=20, max_iters=40, rtol=1e-8,
                                 precond=A.PC_MULTIGRID, sweeps=1)
            ns.newton_step(opts)
        out[mode] = ns.state_host(include_ghosts=False)
        ns.close()
    os.environ.pop("FCFD_MG_NO_LEVEL_LAUNCH", None)
    print("  max |W_level - W_per_colour| = %.3e" % np.abs(out["level"] - out["per_colour"]).max(), flush=True)
    assert np.array_equal(out["level"], out["per_colour"])


def test_wedge_newton_converges():
    from freecfd_amd.ns import NavierStokes
    cs = cases.wedge_case(96, 48)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    opts = A.newton_opts(cfl=10.0, restart=30, max_iters=30)
    hist = []
    for step in range(50):
        st = ns.newton_step(opts)
        hist.append(st.res_norm)
    print("  ||R|| history", " ".join("%.3e" % h for h in hist[::5]), flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.5 * hist[0]


@pytest.mark.parametrize("jac,fp32", [(A.JAC_FD, False), (A.JAC_JVP_FULL, True)])
def test_cylinder_multigrid_newton_reaches_1e10(jac, fp32):
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(64, 32)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    opts = A.newton_opts(cfl=1000.0, jacobian=jac, fp32=fp32, restart=30, max_iters=60, rtol=1e-4, precond=A.PC_MULTIGRID,
                         sweeps=1)
    r = ns.solve(opts, max_steps=80, drop=1e-10)
    print("  cylinder 64x32 jac=%d fp32=%s: %s in %d steps, drop %.2e" % (jac, fp32, r["status"], r["steps"], r["drop"]),
          flush=True)
    assert r["status"] == "converged"


def test_multigrid_graph_replays_match_eager_launches():
    """The V-cycle replayed from its HIP graph (captured once, replayed by
    every GMRES iteration with the eager Krylov kernels in between, over three
    implicit steps that re-assemble the coarse levels) gives bit-identical
    states to the same kernels launched one by one (FCFD_NO_GRAPHS=1)."""
    import os
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(64, 32)
    out = {}
    for mode in ("eager", "graph"):
        if mode == "eager":
            os.environ["FCFD_NO_GRAPHS"] = "1"
        else:
            os.environ.pop("FCFD_NO_GRAPHS", None)
        ns = NavierStokes(cs.raw, cs.opt)
        ns.set_state(cs.W0)
        for fp32, jac in ((False, A.JAC_FD), (True, A.JAC_JVP_FULL), (False, A.JAC_FD)):
            opts = A.newton_opts(cfl=100.0, jacobian=jac, fp32=fp32, restart=20, max_iters=40, rtol=1e-8,
                                 precond=A.PC_MULTIGRID, sweeps=1)
            ns.newton_step(opts)
        out[mode] = ns.state_host(include_ghosts=False)
        ns.close()
    os.environ.pop("FCFD_NO_GRAPHS", None)
    diff = np.abs(out["graph"] - out["eager"]).max()
    print("  max |W_graph - W_eager| = %.3e" % diff, flush=True)
    assert np.array_equal(out["graph"], out["eager"])


def test_relative_update_limit_scales_the_step():
    """max_rel_update scales the whole Newton update by relax < 1 when the
    full step would change p or T by more than the limit: the limited step
    equals relax times the full one, and relax is the limit over the largest
    relative change."""
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(32, 16)
    full = None
    for limit in (0.0, 1e-5):
        # a fresh context each time: the FD step of an implicit step depends on
        # the context's previous update
        ns = NavierStokes(cs.raw, cs.opt)
        ns.set_state(cs.W0)
        opts = A.newton_opts(cfl=50.0, restart=30, max_iters=60, rtol=1e-10, precond=A.PC_BLOCK_SGS, max_rel_update=limit)
        st = ns.newton_step(opts)
        dW = ns.state_host(include_ghosts=False) - cs.W0
        if limit == 0.0:
            full = dW
            assert st.relax == 1.0
            rel = max(np.max(np.abs(dW[:, 0]) / np.abs(cs.W0[:, 0])), np.max(np.abs(dW[:, 4]) / np.abs(cs.W0[:, 4])))
        else:
            print("  relax %.6e, limit/max-rel %.6e" % (st.relax, limit / rel), flush=True)
            assert np.isclose(st.relax, limit / rel, rtol=1e-12)
            assert np.allclose(dW, st.relax * full, rtol=1e-9, atol=1e-12 * np.abs(cs.W0).max())
        ns.close()


@pytest.mark.parametrize("dt_type", [A.DT_FIXED, A.DT_CFL_MAX, A.DT_CFL_LOCAL])
def test_time_step_types_match_the_oracle(dt_type):
    """fcfd_time_step for each update_time_step type (time_step.cc:155-213):
    the cells' dt and the largest CFL number bitwise equal to the CPU
    restatement's on the cylinder (minima and maxima are exact)."""
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = cases.cylinder_case(32, 16, noise=0.05)
    value = 3.0e-6 if dt_type == A.DT_FIXED else 7.5
    o = Oracle(cs.raw, cs.opt)
    o.set_state(cs.W0)
    dt_o, mc_o = o.time_step(dt_type, value)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    dt_g, mc_g = ns.time_step(dt_type, value)
    print("  dt_type %d: max cfl %.17g / %.17g" % (dt_type, mc_g, mc_o), flush=True)
    assert np.array_equal(dt_g, dt_o) and mc_g == mc_o
    ns.close()
    o.close()


def test_newton_step_with_a_global_time_step():
    """A first Newton step with CFL_MAX equals the first step with FIXED at
    the step size CFL_MAX picks (the smallest local step), bitwise; each in a
    fresh context, since a context carries the previous update into the
    next FD Jacobian and initial guess."""
    from freecfd_amd.ns import NavierStokes
    cs = cases.wedge_case(24, 12)
    ns = NavierStokes(cs.raw, cs.opt)
    ns.set_state(cs.W0)
    dt, _ = ns.time_step(A.DT_CFL_MAX, 50.0)
    ns.close()
    out = []
    for dt_type, value in ((A.DT_CFL_MAX, 50.0), (A.DT_FIXED, float(dt[0]))):
        ns = NavierStokes(cs.raw, cs.opt)
        ns.set_state(cs.W0)
        o = A.newton_opts(cfl=value, max_iters=30, rtol=1e-8)
        o.dt_type = dt_type
        st = ns.newton_step(o)
        assert np.isfinite(st.res_norm)
        out.append(ns.state_host(include_ghosts=False))
        ns.close()
    print("  max |W_cflmax - W_fixed| = %.3e, max |dW| = %.3e" % (np.abs(out[0] - out[1]).max(), np.abs(out[0] - cs.W0).max()),
          flush=True)
    assert not np.array_equal(out[0], cs.W0)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("fp32", [False, True])
def test_sliced_matvec_newton_steps_match_the_face_block_matvec(fp32):
    """The assembled operator's matvec streams the signed off-diagonal blocks
    from their sliced-ELL copy; it forms the same products and sums in the
    same order as the matvec over the per-face blocks (FCFD_BSR_MATVEC=1):
    two FD-Jacobian Newton steps (block-Jacobi GMRES) give bit-identical
    states, with FP64 and with FP32 Krylov vectors, on a mesh of hexes,
    prisms and boundary faces (SA plate) and a 2-D O-grid."""
    import os
    from freecfd_amd.ns import NavierStokes
    for cs in (cases.plate_case(20, 10, 4, noise=0.01, sa=True), cases.cylinder_case(48, 24)):
        print("case", cs.name, "fp32", fp32, flush=True)
        out = {}
        for mode in ("face_blocks", "sliced"):
            if mode == "face_blocks":
                os.environ["FCFD_BSR_MATVEC"] = "1"
            else:
                os.environ.pop("FCFD_BSR_MATVEC", None)
            ns = NavierStokes(cs.raw, cs.opt)
            ns.set_state(cs.W0)
            opts = A.newton_opts(cfl=20.0, jacobian=A.JAC_FD, fp32=fp32, restart=20, max_iters=20)
            for _ in range(2):
                ns.newton_step(opts)
            out[mode] = ns.state_host(include_ghosts=False)
            ns.close()
        os.environ.pop("FCFD_BSR_MATVEC", None)
        print("  max |W_sliced - W_face_blocks| = %.3e" % np.abs(out["sliced"] - out["face_blocks"]).max(), flush=True)
        assert np.array_equal(out["sliced"], out["face_blocks"])


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("jac", ["full", "frozen"])
def test_jvp_operator_reading_the_step_gradients_matches_recomputing_them(jac, fp32):
    """Inside a Newton step W is fixed, so the matrix-free operator reads the
    primal gradients, limiters and ghosts that the step's residual wrote
    instead of recomputing them in every product; they are the same numbers,
    so two JVP Newton steps give bit-identical states with and without
    recomputation (FCFD_JVP_RECOMPUTE=1), FP64 and FP32 Krylov, on the SA
    plate (VK limiter, hexes and prisms), the cylinder (unlimited lsqr) and
    the wedge (minmod)."""
    import os
    from freecfd_amd.ns import NavierStokes
    mode = A.JAC_JVP_FULL if jac == "full" else A.JAC_JVP_FROZEN
    for cs in (cases.plate_case(20, 10, 4, noise=0.01, sa=True), cases.cylinder_case(48, 24), cases.wedge_case(24, 12)):
        print("case", cs.name, jac, "fp32", fp32, flush=True)
        out = {}
        for variant in ("recompute", "read"):
            if variant == "recompute":
                os.environ["FCFD_JVP_RECOMPUTE"] = "1"
            else:
                os.environ.pop("FCFD_JVP_RECOMPUTE", None)
            ns = NavierStokes(cs.raw, cs.opt)
            ns.set_state(cs.W0)
            opts = A.newton_opts(cfl=20.0, jacobian=mode, fp32=fp32, restart=20, max_iters=20)
            for _ in range(2):
                ns.newton_step(opts)
            out[variant] = ns.state_host(include_ghosts=False)
            ns.close()
        os.environ.pop("FCFD_JVP_RECOMPUTE", None)
        print("  max |W_read - W_recompute| = %.3e" % np.abs(out["read"] - out["recompute"]).max(), flush=True)
        assert np.array_equal(out["read"], out["recompute"])
