"""Host-side mirror of the reference's NavierStokes/Grid entry points for the
residual path, over the libfcfd C-ABI.

    grid = Grid(raw, opt)                 # Grid::setup .. gradient_maps on the host
    ns = NavierStokes(raw, opt)           # + device upload (fcfd_create)
    ns.set_state(W)                       # p, u, v, w, T per cell (reference order)
    ns.assemble_residual()                # update_boundaries, calc_cell_grads,
                                          # calc_limiter, face loop -> R on the GPU
    R = ns.residual_host()                # N x 5, reference order

Device memory is allocated with torch (plumbing only); every computation runs
in libfcfd's HIP kernels.  Nothing here falls back to the CPU.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from ._lib import lib


def _check(h, rc):
    if rc != 0:
        msg = lib().fcfd_last_error(h)
        raise RuntimeError("libfcfd error %d: %s" % (rc, msg.decode() if msg else ""))


def read_cfl_schedule(path):
    """The (step, CFL) pairs of a CFL schedule file (timemarching
    schedulefile, time_step.cc:63-77): whitespace-separated integers and
    numbers, read in pairs until the end of the file."""
    vals = open(path).read().split()
    return [(int(vals[i]), float(vals[i + 1])) for i in range(0, len(vals) - 1, 2)]


def scheduled_cfl(schedule, step, cfl):
    """update_time_step's schedule (time_step.cc:160-175) at time step `step`
    (from 1): the CFL of the last pair whose step is below `step`, read in
    file order up to the first pair at or beyond it.  Returns (CFL, whether a
    pair applied; the ramp stops once one has).  After the file's last pair
    the CFL stays at its value (the reference's loop reads once more at the
    end of the file; that read changes nothing here)."""
    used = False
    for s, v in schedule:
        if step > s:
            cfl, used = v, True
        else:
            break
    return cfl, used


class Grid:
    """Host-only set-up (no GPU needed): connectivity, metrics, maps."""

    _device = False

    def __init__(self, raw, opt):
        self._mesh = A.make_mesh(raw)
        self._opt = opt
        h = C.c_void_p()
        fn = lib().fcfd_create if self._device else lib().fcfd_create_host
        rc = fn(C.byref(self._mesh), C.byref(opt), C.byref(h))
        if rc != 0:
            raise RuntimeError("fcfd_create failed (%d): %s" % (rc, lib().fcfd_last_error(None).decode()))
        self.h = h
        self.info = A.Info()
        _check(self.h, lib().fcfd_get_info(self.h, C.byref(self.info)))

    def close(self):
        if getattr(self, "h", None):
            lib().fcfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def metrics(self):
        m, arrs = A.alloc_metrics(self.info)
        _check(self.h, lib().fcfd_metrics(self.h, C.byref(m)))
        arrs["face_nodes"] = arrs["face_nodes"][:arrs["face_node_ptr"][self.info.n_faces]]
        return arrs


class NavierStokes(Grid):
    """One grid on one GPU: state, residual, gradients, limiters."""

    _device = True

    def __init__(self, raw, opt, device="cuda:0"):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("NavierStokes needs a GPU (libfcfd has no CPU path)")
        self.torch = torch
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        super().__init__(raw, opt)
        ld, nv = self.info.ld_cell, self.info.nv
        self.W = torch.zeros((nv, ld), dtype=torch.float64, device=self.device)
        self.R = torch.zeros((nv, ld), dtype=torch.float64, device=self.device)

    @property
    def ld(self):
        return self.info.ld_cell

    @property
    def nv(self):
        """variables per cell: 5 (p, u, v, w, T), 6 with the SA model (+ nu~)"""
        return self.info.nv

    def set_state(self, W_host):
        W_host = np.ascontiguousarray(W_host, dtype=np.float64)
        assert W_host.shape == (self.info.n_cells, self.nv)
        self.torch.cuda.synchronize()
        _check(self.h, lib().fcfd_state_to_device(self.h, W_host.ctypes.data, self.W.data_ptr()))

    def assemble_residual(self, W=None, R=None, stream=None):
        """R(W) on the GPU; W, R are [nv][ld] float64 device tensors."""
        W = self.W if W is None else W
        R = self.R if R is None else R
        s = self.torch.cuda.current_stream().cuda_stream if stream is None else stream
        _check(self.h, lib().fcfd_residual(self.h, W.data_ptr(), R.data_ptr(), C.c_void_p(s)))
        return R

    def to_host(self, X, nvar, include_ghosts=False):
        self.torch.cuda.synchronize()
        n = self.info.n_cells + (self.info.n_ghosts if include_ghosts else 0)
        out = np.zeros((n, nvar))
        _check(self.h, lib().fcfd_cells_to_host(self.h, X.data_ptr(), nvar, int(include_ghosts), out.ctypes.data))
        return out

    def residual_host(self):
        return self.to_host(self.R, self.nv)

    def state_host(self, include_ghosts=True):
        return self.to_host(self.W, self.nv, include_ghosts)

    def gradients_limiters(self, W=None, G=None, PHI=None, stream=None):
        """Boundary ghosts, gradients and limiters of W alone on the GPU
        (update_boundaries + calc_cell_grads + calc_limiter) into G and PHI,
        float64 device tensors in row pairs of 2*ceil(3nv/2)*ld and
        2*ceil(nv/2)*ld entries (allocated when not given); returns (G, PHI)."""
        W = self.W if W is None else W
        ld, nv = self.info.ld_cell, self.nv
        if G is None:
            G = self.torch.zeros(2 * ((3 * nv + 1) // 2) * ld, dtype=self.torch.float64, device=self.device)
        if PHI is None:
            PHI = self.torch.zeros(2 * ((nv + 1) // 2) * ld, dtype=self.torch.float64, device=self.device)
        s = self.torch.cuda.current_stream().cuda_stream if stream is None else stream
        _check(self.h, lib().fcfd_gradients_limiters(self.h, W.data_ptr(), G.data_ptr(), PHI.data_ptr(), C.c_void_p(s)))
        return G, PHI

    def gradients_limiters_host(self):
        """Gradients ((N+B) x 3nv) and limiters (N x nv) of the last residual."""
        n, nv = self.info.n_cells, self.nv
        G = np.zeros((n + self.info.n_ghosts, 3 * nv))
        PHI = np.zeros((n, nv))
        _check(self.h, lib().fcfd_gradients_to_host(self.h, G.ctypes.data, PHI.ctypes.data))
        return G, PHI

    def fd_jacobian(self, W=None, dW_prev=None, dt=None, stream=None):
        """Reference FD Jacobian (+ time terms if dt is given) on the GPU."""
        W = self.W if W is None else W
        s = self.torch.cuda.current_stream().cuda_stream if stream is None else stream
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        _check(self.h, lib().fcfd_fd_jacobian(self.h, W.data_ptr(), ptr(dW_prev), ptr(dt), C.c_void_p(s)))

    def jacobian_host(self):
        F, N, nv = self.info.n_faces, self.info.n_cells, self.nv
        JL = np.zeros((F, nv, nv))
        JR = np.zeros((F, nv, nv))
        D = np.zeros((N, nv, nv))
        _check(self.h, lib().fcfd_jacobian_to_host(self.h, JL.ctypes.data, JR.ctypes.data, D.ctypes.data))
        return JL, JR, D

    def jvp(self, dW, mode=A.JAC_JVP_FULL, W=None, out=None):
        """dR = (dR/dW) dW on the GPU (dW a [nv][ld] float64 device tensor)."""
        W = self.W if W is None else W
        out = self.torch.zeros_like(self.W) if out is None else out
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_jvp(self.h, W.data_ptr(), dW.data_ptr(), out.data_ptr(), int(mode), C.c_void_p(s)))
        return out

    def convective_flux(self, flux, WL, WR, normal, tangent1, area, diss=None, wall_dist=None):
        """The convective flux of a batch of faces from given states
        (fcfd_convective_flux: roe_flux and its siblings with
        convective_face_flux's rotation and area).  WL, WR: (5, n) p, u, v,
        w, T; normal, tangent1: (3, n); area, diss, wall_dist: (n,); host
        arrays or device tensors.  Returns (flux (5, n), weightL (n,)) as
        device tensors."""
        torch = self.torch
        dev = lambda x: None if x is None else torch.as_tensor(np.asarray(x, dtype=np.float64) if not torch.is_tensor(x) else x,
                                                                dtype=torch.float64, device=self.device).contiguous()
        WL, WR, normal, tangent1, area = dev(WL), dev(WR), dev(normal), dev(tangent1), dev(area)
        diss, wall_dist = dev(diss), dev(wall_dist)
        n = int(area.numel())
        if WL.shape != (5, n) or WR.shape != (5, n) or normal.shape != (3, n) or tangent1.shape != (3, n):
            raise ValueError("convective_flux: WL/WR must be (5, n) and normal/tangent1 (3, n) with n = len(area)")
        for x in (diss, wall_dist):
            if x is not None and x.numel() != n:
                raise ValueError("convective_flux: diss and wall_dist must have n entries")
        out = torch.empty((5, n), dtype=torch.float64, device=self.device)
        wl = torch.empty((n,), dtype=torch.float64, device=self.device)
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        s = torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_convective_flux(self.h, int(flux), n, ptr(WL), ptr(WR), ptr(normal), ptr(tangent1), ptr(area),
                                                  ptr(diss), ptr(wall_dist), ptr(out), ptr(wl), C.c_void_p(s)))
        return out, wl

    def vjp(self, Rbar, fx_weight=None, W=None, out=None):
        """Wbar = (dR/dW)^T Rbar [+ d/dW sum_f fx_weight[bc(f)] . (D - C)_f] on
        the GPU; Rbar a [5][ld] float64 device tensor, fx_weight an optional
        (n_bc, 5) array (host or device)."""
        W = self.W if W is None else W
        out = self.torch.zeros_like(self.W) if out is None else out
        fw = None
        if fx_weight is not None:
            fw = self.torch.as_tensor(np.asarray(fx_weight, dtype=np.float64).reshape(-1), device=self.W.device).contiguous()
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_vjp(self.h, W.data_ptr(), Rbar.data_ptr(), None if fw is None else fw.data_ptr(),
                                      out.data_ptr(), C.c_void_p(s)))
        return out

    def bc_tangent(self, index, W=None, out=None, **direction):
        """Exact derivative at the fixed state W of the residual and of the
        loads with respect to the values of BC `index`, in the direction
        given by keywords (p, T, rho, V, p_total, T_total, mdot, qdot, nut):
        (dR [nv][ld] device tensor, dloads (n_bc, 6))."""
        W = self.W if W is None else W
        out = self.torch.zeros_like(self.W) if out is None else out
        d = A.BC()
        for k, v in direction.items():
            if k == "V":
                for i in range(3):
                    d.V[i] = float(v[i])
            else:
                setattr(d, k, float(v))
        n_bc = max(int(self._opt.n_bc), 1)
        dl = np.zeros(6 * n_bc)
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_bc_tangent(self.h, W.data_ptr(), int(index), C.byref(d), out.data_ptr(), dl.ctypes.data,
                                             C.c_void_p(s)))
        return out, dl.reshape(n_bc, 6)[: int(self._opt.n_bc)]

    def vjp_params(self, Rbar, fx_weight=None, W=None):
        """pbar (n_bc, 11): <Rbar, dR/dv> + d/dv sum_f fx_weight[bc(f)] . (D - C)_f
        for every BC value v (A.BC_PARAMS order), exact (fcfd_vjp_params)."""
        W = self.W if W is None else W
        n_bc = int(self._opt.n_bc)
        out = np.zeros(len(A.BC_PARAMS) * max(n_bc, 1))
        fw = None
        if fx_weight is not None:
            fw = self.torch.as_tensor(np.asarray(fx_weight, dtype=np.float64).reshape(-1), device=self.W.device).contiguous()
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_vjp_params(self.h, W.data_ptr(), Rbar.data_ptr(), None if fw is None else fw.data_ptr(),
                                             out.ctypes.data, C.c_void_p(s)))
        return out.reshape(max(n_bc, 1), len(A.BC_PARAMS))[:n_bc]

    def loads(self, W=None):
        """(n_bc, 6) surface loads (Fx Fy Fz Mx My Mz) of the state, reference
        face-order sums over the BCs with include_in_loads."""
        W = self.W if W is None else W
        n_bc = max(int(self._opt.n_bc), 1)
        out = np.zeros(6 * n_bc)
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_loads(self.h, W.data_ptr(), out.ctypes.data, C.c_void_p(s)))
        return out.reshape(n_bc, 6)[: int(self._opt.n_bc)]

    def pseudo_step(self, opts, dtau_type, dtau, ps_step, Wn=None, ws95=False, W=None, stream=None, check=True):
        """One pseudo step of dual time stepping (fcfd_pseudo_step): opts sets
        the physical step and the Krylov solve, (dtau_type, dtau) the pseudo
        step, Wn ([nv][ld] device tensor) the state at the time level."""
        W = self.W if W is None else W
        ps = A.PseudoOpts(dtau_type=dtau_type, dtau=dtau, preconditioner=A.PSEUDO_PC_WS95 if ws95 else A.PSEUDO_PC_NONE,
                          ps_step=ps_step, Wn_dev=None if Wn is None else Wn.data_ptr())
        st = A.NewtonStats()
        s = self.torch.cuda.current_stream().cuda_stream if stream is None else stream
        rc = lib().fcfd_pseudo_step(self.h, W.data_ptr(), C.byref(opts), C.byref(ps), C.byref(st), C.c_void_p(s))
        if check:
            _check(self.h, rc)
            return st
        return rc, st

    def surface(self, W=None):
        """Surface quantities of the face loop (ns_assemble_linear_system.cc:
        122-127) in reference face order: (mdot[F], qdot[F], tau[F, 3]);
        qdot and tau are set on boundary faces, 0 elsewhere."""
        W = self.W if W is None else W
        F = self.info.n_faces
        mdot, qdot, tau = np.zeros(F), np.zeros(F), np.zeros(3 * F)
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_surface(self.h, W.data_ptr(), mdot.ctypes.data, qdot.ctypes.data, tau.ctypes.data,
                                          C.c_void_p(s)))
        return mdot, qdot, tau.reshape(F, 3)

    def to_device(self, X_host):
        """N x nv host array (reference order) -> [nv][ld] device tensor."""
        X_host = np.ascontiguousarray(X_host, dtype=np.float64)
        t = self.torch.zeros_like(self.W)
        self.torch.cuda.synchronize()
        _check(self.h, lib().fcfd_state_to_device(self.h, X_host.ctypes.data, t.data_ptr()))
        return t

    def newton_step(self, opts=None, W=None, stream=None, check=True):
        """One implicit step (NavierStokes::solve, ns.cc:104-118); returns NewtonStats."""
        opts = A.newton_opts() if opts is None else opts
        W = self.W if W is None else W
        st = A.NewtonStats()
        s = self.torch.cuda.current_stream().cuda_stream if stream is None else stream
        rc = lib().fcfd_newton_step(self.h, W.data_ptr(), C.byref(opts), C.byref(st), C.c_void_p(s))
        if check:
            _check(self.h, rc)
        return st if check else (rc, st)

    def solve(self, opts=None, max_steps=500, drop=1e-10, cfl_initial=None, cfl_growth=1.0, cfl_target=None,
              W=None, callback=None, schedule=None):
        """Implicit pseudo-time iterations to a residual drop (the step loop of
        main.cc/NavierStokes::solve with update_time_step's CFL ramp,
        time_step.cc:143-158: CFL_1 = cfl_initial, CFL_k = min(CFL_{k-1} *
        growth, target)).  schedule: the (step, CFL) pairs of a CFL schedule
        file (read_cfl_schedule); the step type is then CFL_MAX and the
        schedule takes over from the ramp (time_step.cc:160-175).  Stops when
        ||R_k||_2 <= drop * ||R_0||_2 (R before step k), on divergence, or
        after max_steps.  Returns a dict with the residual history, per-step
        wall times (ms) and the final drop."""
        import time as _time
        opts = A.newton_opts() if opts is None else opts
        cfl = opts.cfl if cfl_initial is None else cfl_initial
        target = cfl if cfl_target is None else cfl_target
        ramp = True
        if schedule is not None:
            opts.dt_type = A.DT_CFL_MAX
        hist, ms, iters = [], [], []
        status = "max_steps"
        for k in range(max_steps):
            if ramp and k > 0:
                cfl = min(cfl * cfl_growth, target)
            if schedule is not None:
                c, used = scheduled_cfl(schedule, k + 1, cfl)
                if used:
                    cfl, ramp = c, False
            opts.cfl = cfl
            self.torch.cuda.synchronize()
            t0 = _time.perf_counter()
            rc, st = self.newton_step(opts, W=W, check=False)
            self.torch.cuda.synchronize()
            ms.append(1e3 * (_time.perf_counter() - t0))
            hist.append(st.res_norm)
            iters.append(st.krylov_iters)
            if callback is not None:
                callback(k, st, cfl)
            if rc != 0 or not np.isfinite(st.res_norm):
                status = "diverged"
                break
            if st.res_norm <= drop * hist[0]:
                status = "converged"
                break
        return dict(status=status, res_norm=hist, ms=ms, krylov_iters=iters, steps=len(hist),
                    drop=hist[-1] / hist[0] if hist else 1.0, total_ms=float(np.sum(ms)))

    def set_bc(self, index, **values):
        """Replace values of BC `index` (p, T, rho, V, p_total, T_total, mdot,
        qdot, nut) in the live context; the mesh set-up is unchanged."""
        bc = A.BC.from_buffer_copy(self._opt.bc[index])
        for k, v in values.items():
            if k == "V":
                for i in range(3):
                    bc.V[i] = float(v[i])
            else:
                setattr(bc, k, float(v))
        self.torch.cuda.synchronize()
        _check(self.h, lib().fcfd_set_bc(self.h, int(index), C.byref(bc)))
        self._opt.bc[index] = bc

    def adjoint_step(self, g, lam, opts=None, W=None):
        """One pseudo-time step of (dR/dW)^T lam = -g (g, lam [nv][ld] device
        tensors, lam updated in place); returns NewtonStats with res_norm =
        ||(dR/dW)^T lam + g||_2 before the step."""
        opts = A.newton_opts(fp32=False) if opts is None else opts
        W = self.W if W is None else W
        st = A.NewtonStats()
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_adjoint_step(self.h, W.data_ptr(), g.data_ptr(), lam.data_ptr(), C.byref(opts), C.byref(st),
                                               C.c_void_p(s)))
        return st

    def adjoint_solve(self, g, opts=None, max_steps=200, drop=1e-10, cfl_initial=None, cfl_growth=1.0, cfl_target=None,
                      lam=None, W=None):
        """Adjoint pseudo-time iterations from lam = 0 (or `lam`) with the same
        CFL ramp as solve(), until ||(dR/dW)^T lam + g|| <= drop * ||g||.
        Returns (lam, dict(status, res_norm history, steps, drop))."""
        opts = A.newton_opts(fp32=False) if opts is None else opts
        lam = self.torch.zeros_like(self.W) if lam is None else lam
        cfl = opts.cfl if cfl_initial is None else cfl_initial
        target = cfl if cfl_target is None else cfl_target
        g0 = float(self.torch.linalg.vector_norm(g))
        hist, status = [], "max_steps"
        for k in range(max_steps):
            if k > 0:
                cfl = min(cfl * cfl_growth, target)
            opts.cfl = cfl
            st = self.adjoint_step(g, lam, opts, W=W)
            hist.append(st.res_norm)
            if not np.isfinite(st.res_norm):
                status = "diverged"
                break
            if st.res_norm <= drop * g0:
                status = "converged"
                break
        return lam, dict(status=status, res_norm=hist, steps=len(hist), drop=(hist[-1] / g0) if hist and g0 > 0 else 0.0)

    def time_step_host(self, cfl):
        dt = self.torch.zeros(self.ld, dtype=self.torch.float64, device=self.device)
        mc = C.c_double()
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_time_step(self.h, self.W.data_ptr(), A.DT_CFL_LOCAL, cfl, dt.data_ptr(), C.byref(mc),
                                           C.c_void_p(s)))
        return self.to_host(dt.reshape(1, -1), 1)[:, 0]

    def time_step(self, dt_type, value):
        """update_time_step for the current state (fcfd_time_step):
        FCFD_DT_FIXED (value = the step size), FCFD_DT_CFL_MAX or
        FCFD_DT_CFL_LOCAL (value = the CFL number).  Returns the cells' dt in
        reference order and the largest CFL number."""
        dt = self.torch.zeros(self.ld, dtype=self.torch.float64, device=self.device)
        mc = C.c_double()
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_time_step(self.h, self.W.data_ptr(), int(dt_type), float(value), dt.data_ptr(), C.byref(mc),
                                           C.c_void_p(s)))
        return self.to_host(dt.reshape(1, -1), 1)[:, 0], mc.value

    def status(self, clear=True):
        _check(self.h, lib().fcfd_status(self.h, int(clear)))

    # ---- segregated two-equation RANS (the reference's RANS object, src/rans/) ----
    def rans_initialize(self, intensity, visc_ratio, opts=None, W=None):
        """RANS::initialize (rans.cc:70-81) at the NS state: initial k,
        omega over the domain from the turbulence intensity and eddy
        viscosity ratio, the inlets' values, ghosts, mu_t, gradients.  opts
        (RansOpts) are used by the solve inside newton_step."""
        W = self.W if W is None else W
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_rans_init(self.h, W.data_ptr(), float(intensity), float(visc_ratio),
                                            None if opts is None else C.byref(opts), C.c_void_p(s)))

    def rans_set(self, K, W=None):
        """k, omega of the interior cells (N x 2, reference order), then
        ghosts, mu_t and gradients"""
        W = self.W if W is None else W
        K = np.ascontiguousarray(K, dtype=np.float64)
        assert K.shape == (self.info.n_cells, 2)
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_rans_set(self.h, W.data_ptr(), K.ctypes.data, C.c_void_p(s)))

    def rans_fields(self):
        """dict in reference order: K (N+B) x 2, mut N+B, G N x 6 (grad k,
        grad omega), strain N+B, mutf F (mu_t.face)"""
        N, B, F = self.info.n_cells, self.info.n_ghosts, self.info.n_faces
        d = dict(K=np.zeros((N + B, 2)), mut=np.zeros(N + B), G=np.zeros((N, 6)), strain=np.zeros(N + B), mutf=np.zeros(F))
        _check(self.h, lib().fcfd_rans_get(self.h, d["K"].ctypes.data, d["mut"].ctypes.data, d["G"].ctypes.data,
                                           d["strain"].ctypes.data, d["mutf"].ctypes.data))
        return d

    def rans_terms(self, dt, W=None):
        """RANS::terms + time_terms with the NS time step dt ([ld] device
        tensor): (rhs N x 2, diag N x 2 x 2, jf F x 6) in reference order"""
        W = self.W if W is None else W
        N, F = self.info.n_cells, self.info.n_faces
        rhs, diag, jf = np.zeros((N, 2)), np.zeros((N, 2, 2)), np.zeros((F, 6))
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_rans_terms(self.h, W.data_ptr(), dt.data_ptr(), rhs.ctypes.data, diag.ctypes.data,
                                             jf.ctypes.data, C.c_void_p(s)))
        return rhs, diag, jf

    def rans_solve(self, dt, opts=None, W=None, check=True):
        """RANS::solve (rans.cc:85-98) at the NS state with the NS time step
        dt ([ld] device tensor); returns RansStats"""
        W = self.W if W is None else W
        opts = A.rans_opts() if opts is None else opts
        st = A.RansStats()
        s = self.torch.cuda.current_stream().cuda_stream
        rc = lib().fcfd_rans_step(self.h, W.data_ptr(), dt.data_ptr(), C.byref(opts), C.byref(st), C.c_void_p(s))
        if check:
            _check(self.h, rc)
            return st
        return rc, st

    def time_step_dev(self, dt_type, value, W=None):
        """fcfd_time_step into a fresh [ld] device tensor"""
        W = self.W if W is None else W
        dt = self.torch.zeros(self.ld, dtype=self.torch.float64, device=self.device)
        mc = C.c_double()
        s = self.torch.cuda.current_stream().cuda_stream
        _check(self.h, lib().fcfd_time_step(self.h, W.data_ptr(), int(dt_type), float(value), dt.data_ptr(), C.byref(mc),
                                           C.c_void_p(s)))
        return dt
