"""The CPU side of the Newton comparison (the restatement's implicit step on
the host cores): its block-Jacobi GMRES solves the FD-Jacobian system of the
step (checked against a direct sparse solve of the same blocks), its result
does not depend on the thread count, and steps reduce the residual."""
import os
import subprocess
import sys

import numpy as np
import pytest

from freecfd_amd import cases


def _oracle(cs):
    from oracle.oracle import Oracle
    return Oracle(cs.raw, cs.opt)


def test_cpu_newton_step_solves_the_fd_system(oracle_lib):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    cs = cases.cylinder_case(24, 12, noise=0.01)
    o = _oracle(cs)
    o.set_state(cs.W0)
    st = o.newton_step(cfl=50.0, restart=60, max_iters=600, rtol=1e-12)
    W1 = o.state()
    dW = W1 - cs.W0
    # the same system assembled independently: M = diag(D + P V/dt) - offdiag
    o2 = _oracle(cs)
    R, _ = o2.residual(cs.W0, with_jacobian=True)
    JL, JR, D = o2.jacobian()
    dt, _ = o2.time_step(3, 50.0)
    T = o2.time_terms(dt)
    m = o2.metrics()
    N, nv = cs.n_cells, 5
    rows, cols, vals = [], [], []
    def add(r, c, B):
        rr, cc = np.meshgrid(np.arange(nv), np.arange(nv), indexing="ij")
        rows.append((nv * r + rr).ravel()); cols.append((nv * c + cc).ravel()); vals.append(B.ravel())
    for c in range(N):
        add(c, c, D[c] + T[c])
    for f in range(o2.info.n_faces):
        p, n = m["face_parent"][f], m["face_neighbor"][f]
        if m["face_bc"][f] < 0:
            add(p, n, -JR[f].T)
            add(n, p, JL[f].T)
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nv * N, nv * N))
    x = spla.spsolve(M.tocsc(), R.ravel()).reshape(N, nv)
    err = np.abs(dW - x).max(axis=0) / np.abs(x).max(axis=0)
    print("  CPU GMRES vs direct solve: %s, %d iterations" % (" ".join("%.1e" % e for e in err), st["krylov_iters"]))
    assert err.max() < 1e-8
    assert np.isclose(st["res_norm"], np.linalg.norm(R), rtol=1e-14)


def test_cpu_newton_step_is_thread_independent(tmp_path, oracle_lib):
    code = r'''
import sys, numpy as np
sys.path.insert(0, %r)
from freecfd_amd import cases
from oracle.oracle import Oracle
cs = cases.wedge_case(16, 8)
cs.opt.limiter = -1     # unlimited: a limiter switch would amplify the reductions' rounding
o = Oracle(cs.raw, cs.opt)
o.set_state(cs.W0)
for _ in range(3):
    o.newton_step(cfl=10.0, restart=20, max_iters=20, rtol=1e-8)
np.save(sys.argv[1], o.state())
''' % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = {}
    for nt in (1, 4):
        env = dict(os.environ, OMP_NUM_THREADS=str(nt))
        path = str(tmp_path / ("w%d.npy" % nt))
        subprocess.check_call([sys.executable, "-c", code, path], env=env)
        out[nt] = np.load(path)
    # the residual and Jacobian are bitwise independent of the thread count;
    # the GMRES dot products are OpenMP reductions, so only the Krylov
    # rounding may differ
    # (w is rounding noise on the one-cell-thick wedge: velocities are
    # compared against the speed)
    a, b = out[1], out[4]
    speed = np.abs(a[:, 1:4]).max()
    scale = np.array([np.abs(a[:, 0]).max(), speed, speed, speed, np.abs(a[:, 4]).max()])
    assert np.all(np.abs(a - b) <= 1e-9 * scale)


def test_cpu_newton_steps_reduce_the_residual(oracle_lib):
    cs = cases.wedge_case(24, 12)
    o = _oracle(cs)
    o.set_state(cs.W0)
    hist = [o.newton_step(cfl=10.0, restart=30, max_iters=30)["res_norm"] for _ in range(8)]
    print("  ||R||:", " ".join("%.3e" % h for h in hist))
    assert hist[-1] < 0.5 * hist[0]
