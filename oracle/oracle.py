"""ctypes wrapper of the CPU restatement (TEST INFRASTRUCTURE ONLY).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the parity checker; the product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from freecfd_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(A.Mesh), C.POINTER(A.Options)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [C.c_void_p]
        for name in ("orc_set_state", "orc_set_update", "orc_set_full_state"):
            getattr(L, name).argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        for name in ("orc_update_boundaries", "orc_calc_cell_grads", "orc_calc_limiter"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_assemble.argtypes = [C.c_void_p, C.c_int]
        L.orc_info.argtypes = [C.c_void_p, C.POINTER(A.Info)]
        L.orc_metrics.argtypes = [C.c_void_p, C.POINTER(A.MetricsOut)]
        L.orc_get_residual.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_fields.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_jacobian.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_loads.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_time_step.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.POINTER(C.c_double)]
        L.orc_time_terms.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_flux_probe.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_void_p]
        L.orc_sa_probe.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        L.orc_get_face_flux.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_surface.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_newton_step.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double, C.c_void_p]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_wall_brute_force.argtypes = [C.c_int]
        L.orc_pseudo_system.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        L.orc_bc_probe.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 6
        L.orc_ausm_probe.argtypes = [C.c_double, C.c_double, C.c_void_p]
        L.orc_gelimd_probe.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_wtli_probe.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        L.orc_kw_init.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_kw_set.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_kw_bc_values.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_kw_terms.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_kw_update.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_kw_get.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.orc_get_weight_left.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_kw_source_probe.argtypes = [C.c_int] + [C.c_double] * 5 + [C.c_void_p] * 3 + [C.c_double] * 2 + [C.c_void_p]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Reference-numbered CPU restatement of one grid + NavierStokes solver."""

    def __init__(self, raw, opt):
        L = lib()
        self._mesh = A.make_mesh(raw)
        self._opt = opt
        h = L.orc_create(C.byref(self._mesh), C.byref(opt))
        if not h:
            raise RuntimeError("oracle set-up failed: " + L.orc_last_error(None).decode())
        self.h = h
        self.info = A.Info()
        L.orc_info(h, C.byref(self.info))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError("oracle error %d: %s" % (rc, lib().orc_last_error(self.h).decode()))

    def metrics(self):
        m, arrs = A.alloc_metrics(self.info)
        self._check(lib().orc_metrics(self.h, C.byref(m)))
        F = self.info.n_faces
        arrs["face_nodes"] = arrs["face_nodes"][:arrs["face_node_ptr"][F]]
        return arrs

    def set_state(self, W, dW=None):
        W = np.ascontiguousarray(W, dtype=np.float64)
        self._check(lib().orc_set_state(self.h, W.ctypes.data_as(C.POINTER(C.c_double))))
        if dW is not None:
            dW = np.ascontiguousarray(dW, dtype=np.float64)
            self._check(lib().orc_set_update(self.h, dW.ctypes.data_as(C.POINTER(C.c_double))))

    def set_full_state(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64)
        self._check(lib().orc_set_full_state(self.h, W.ctypes.data_as(C.POINTER(C.c_double))))

    def grads(self):
        self._check(lib().orc_calc_cell_grads(self.h))
        return self.fields()[1]

    def limiter(self):
        self._check(lib().orc_calc_limiter(self.h))
        return self.fields()[2]

    def flux_probe(self, flux_type, WL, WR, n, t1, area=1.0, weight=False):
        """convective_face_flux of one face: the flux (5,), and with weight
        also weightL"""
        out = np.zeros(6)
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (WL, WR, n, t1)]
        self._check(lib().orc_flux_probe(self.h, flux_type, _p(arr[0]), _p(arr[1]), _p(arr[2]), _p(arr[3]),
                                         float(area), _p(out)))
        return (out[:5], out[5]) if weight else out[:5]

    def prepare(self):
        """update_boundaries -> calc_cell_grads -> calc_limiter (ns.cc:112-116)."""
        L = lib()
        self._check(L.orc_update_boundaries(self.h))
        self._check(L.orc_calc_cell_grads(self.h))
        self._check(L.orc_calc_limiter(self.h))

    def residual(self, W, with_jacobian=False, dW=None):
        self.set_state(W, dW)
        self.prepare()
        self._check(lib().orc_assemble(self.h, int(with_jacobian)))
        N, nv = self.info.n_cells, self.info.nv
        R = np.zeros((N, nv))
        Rabs = np.zeros((N, nv))
        self._check(lib().orc_get_residual(self.h, _p(R), _p(Rabs)))
        return R, Rabs

    def fields(self):
        T = self.info.n_cells + self.info.n_ghosts
        nv = self.info.nv
        W = np.zeros((T, nv))
        G = np.zeros((T, 3 * nv))
        phi = np.zeros((self.info.n_cells, nv))
        self._check(lib().orc_get_fields(self.h, _p(W), _p(G), _p(phi)))
        return W, G, phi

    def jacobian(self):
        F, N, nv = self.info.n_faces, self.info.n_cells, self.info.nv
        JL = np.zeros((F, nv, nv))
        JR = np.zeros((F, nv, nv))
        Dg = np.zeros((N, nv, nv))
        self._check(lib().orc_get_jacobian(self.h, _p(JL), _p(JR), _p(Dg)))
        return JL, JR, Dg

    def face_fluxes(self):
        """(C, D): convective and diffusive flux of every face of the last
        residual evaluation, F x nv each, reference face order."""
        F, nv = self.info.n_faces, self.info.nv
        Cf = np.zeros((F, nv))
        Df = np.zeros((F, nv))
        self._check(lib().orc_get_face_flux(self.h, _p(Cf), _p(Df)))
        return Cf, Df

    def newton_step(self, cfl, restart=30, max_iters=30, rtol=1e-6):
        """one implicit step on the host cores from the current state (set
        with set_state); returns dict(res_norm, krylov_iters, ms_residual_jacobian,
        ms_precond, ms_krylov, ms_update)"""
        st = np.zeros(6)
        self._check(lib().orc_newton_step(self.h, float(cfl), int(restart), int(max_iters), float(rtol), _p(st)))
        return dict(res_norm=st[0], krylov_iters=int(st[1]), ms_residual_jacobian=st[2], ms_precond=st[3], ms_krylov=st[4],
                    ms_update=st[5])

    def state(self):
        """interior state N x nv"""
        return self.fields()[0][:self.info.n_cells].copy()

    def pseudo_system(self, dt, dtau, ws95, ps_step, Wn):
        """after residual(): (blocks N x nv x nv = P V/dt + P_ps V/dtau,
        rhs N x nv = R - (P V/dt)(W - Wn) for ps_step > 1)"""
        N, nv = self.info.n_cells, self.info.nv
        blocks, rhs = np.zeros((N, nv, nv)), np.zeros((N, nv))
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (dt, dtau, Wn)]
        self._check(lib().orc_pseudo_system(self.h, _p(arr[0]), _p(arr[1]), int(ws95), int(ps_step), _p(arr[2]), _p(blocks),
                                            _p(rhs)))
        return blocks, rhs

    def surface(self):
        """(mdot[F], qdot[F], tau[F, 3]) of the last residual evaluation"""
        F = self.info.n_faces
        mdot, qdot, tau = np.zeros(F), np.zeros(F), np.zeros((F, 3))
        self._check(lib().orc_get_surface(self.h, _p(mdot), _p(qdot), _p(tau)))
        return mdot, qdot, tau

    def bc_probe(self, bc, WL, WLc, WRc, n, t1):
        """apply_bcs of BC `bc` on one face: returns (status, dict of the
        right state: p, V, T, rho, a (as the BC set it), p_c, V_c, T_c)."""
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (WL, WLc, WRc, n, t1)]
        out = np.zeros(12)
        rc = lib().orc_bc_probe(self.h, int(bc), *[_p(a) for a in arr], _p(out))
        return rc, dict(p=out[0], V=out[1:4].copy(), T=out[4], rho=out[5], a=out[6], p_c=out[7], V_c=out[8:11].copy(),
                        T_c=out[11])

    def loads(self):
        out = np.zeros((self._opt.n_bc, 6))
        self._check(lib().orc_get_loads(self.h, _p(out)))
        return out

    def time_step(self, dt_type, cfl):
        dt = np.zeros(self.info.n_cells)
        mc = C.c_double(0.0)
        self._check(lib().orc_time_step(self.h, dt_type, cfl, _p(dt), C.byref(mc)))
        return dt, mc.value

    def time_terms(self, dt):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        nv = self.info.nv
        out = np.zeros((self.info.n_cells, nv, nv))
        self._check(lib().orc_time_terms(self.h, _p(dt), _p(out)))
        return out

    # ---- segregated two-equation RANS (src/rans/) ----
    def kw_init(self, intensity=1e-4, visc_ratio=0.1):
        """apply_initial_conditions over the whole domain + set_bcs +
        ghosts, mu_t and gradients, at the current NS state (after
        residual() or prepare())"""
        self._check(lib().orc_kw_init(self.h, float(intensity), float(visc_ratio)))

    def kw_set(self, K):
        K = np.ascontiguousarray(K, dtype=np.float64)
        self._check(lib().orc_kw_set(self.h, _p(K)))

    def kw_bc_values(self, b):
        out = np.zeros(3)
        self._check(lib().orc_kw_bc_values(self.h, int(b), _p(out)))
        return out

    def kw_terms(self, dt):
        """after residual() at the same NS state: (rhs N x 2, diag N x 2 x 2,
        jf F x 6 = parent-side flux contribution k, omega, jacL_k, jacL_omega,
        jacR_k, jacR_omega)"""
        N, F = self.info.n_cells, self.info.n_faces
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        rhs, diag, jf = np.zeros((N, 2)), np.zeros((N, 2, 2)), np.zeros((F, 6))
        self._check(lib().orc_kw_terms(self.h, _p(dt), _p(rhs), _p(diag), _p(jf)))
        return rhs, diag, jf

    def kw_update(self, dU, dt):
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        sums = np.zeros(2)
        self._check(lib().orc_kw_update(self.h, _p(dU), _p(dt), _p(sums)))
        return sums

    def kw_fields(self):
        """dict: K (N+B) x 2, mut (N+B), G (N+B) x 6 (grad k, grad omega),
        strain (N+B), mutf F"""
        T, F = self.info.n_cells + self.info.n_ghosts, self.info.n_faces
        d = dict(K=np.zeros((T, 2)), mut=np.zeros(T), G=np.zeros((T, 6)), strain=np.zeros(T), mutf=np.zeros(F))
        self._check(lib().orc_kw_get(self.h, _p(d["K"]), _p(d["mut"]), _p(d["G"]), _p(d["strain"]), _p(d["mutf"])))
        return d

    def weight_left(self):
        wl = np.zeros(self.info.n_faces)
        self._check(lib().orc_get_weight_left(self.h, _p(wl)))
        return wl


def kw_source_probe(model, rho, k, omega, mu_t, lam_visc, grad_vel, gk, go, d, vol=1.0):
    """one cell of the two-equation source loop: dict(src (2), jac (k,k),
    (k,omega), (omega,omega), F1, strain)"""
    gv = np.ascontiguousarray(grad_vel, dtype=np.float64).reshape(9)
    a = np.ascontiguousarray(gk, dtype=np.float64).reshape(3)
    b = np.ascontiguousarray(go, dtype=np.float64).reshape(3)
    out = np.zeros(7)
    lib().orc_kw_source_probe(int(model), float(rho), float(k), float(omega), float(mu_t), float(lam_visc), _p(gv), _p(a),
                              _p(b), float(d), float(vol), _p(out))
    return dict(src=out[0:2].copy(), jac=out[2:5].copy(), F1=out[5], strain=out[6])


def sa_probe(rho, mu, nut, grad_vel, grad_nut, d):
    """SA model pieces at one point: grad_vel is the 3x3 velocity gradient
    (rows du, dv, dw), grad_nut the nu~ gradient.  Returns a dict with the
    source per unit volume, P, D (per unit mass), mu_t and fn."""
    gv = np.ascontiguousarray(grad_vel, dtype=np.float64).reshape(9)
    gn = np.ascontiguousarray(grad_nut, dtype=np.float64).reshape(3)
    out = np.zeros(5)
    lib().orc_sa_probe(float(rho), float(mu), float(nut), _p(gv), _p(gn), float(d), _p(out))
    return dict(source=out[0], P=out[1], D=out[2], mu_t=out[3], fn=out[4])


def ausm_split(M, alpha):
    """AUSM+-up split functions (M4+, M4-, P5+, P5-) at Mach number M."""
    out = np.zeros(4)
    lib().orc_ausm_probe(float(M), float(alpha), _p(out))
    return out


def gelimd(a, b, x0=None):
    """the reference's Gaussian elimination on a small dense system: returns
    (status, x); x keeps x0 (default zeros) where the system is singular"""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b)) if x0 is None else np.array(x0, dtype=np.float64)
    rc = lib().orc_gelimd_probe(len(b), _p(a), _p(b), _p(x))
    return rc, x


def wtli(point, ids, pts, dim=3, max_stencil=12, tol=0.99):
    """Interpolate::calculate_weights(is_internal = true) on one stencil:
    returns (kind, kept ids, weights); kind 4 tetra, 3 triangle, 2 line, 1 point"""
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    pt = np.ascontiguousarray(point, dtype=np.float64)
    oi = np.zeros(len(ids) + 1, np.int32)
    ow = np.zeros(len(ids) + 1)
    kind = C.c_int(0)
    nw = lib().orc_wtli_probe(int(dim), int(max_stencil), float(tol), _p(pt), len(ids), _p(ids), _p(pts), _p(oi), _p(ow),
                              C.byref(kind))
    return kind.value, oi[:nw].copy(), ow[:nw].copy()


def num_threads():
    """threads the restatement's OpenMP loops use (OMP_NUM_THREADS or all cores)"""
    return int(lib().orc_num_threads())


def set_wall_brute_force(on):
    """wall distances of later Oracle() set-ups by the reference's scan (on)
    or the exact bucketed search (off, the default; bitwise the same)"""
    lib().orc_set_wall_brute_force(int(bool(on)))
