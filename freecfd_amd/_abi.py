"""ctypes mirror of include/fcfd_gpu.h (plain C structs and constants)."""
import ctypes as C

import numpy as np

OK, EINVAL, EMESH, EBC, EHIP, EDIVERGED, ENOTSUP, ENOMEM = 0, -1, -2, -3, -4, -5, -6, -7

BC_SYMMETRY, BC_WALL, BC_INLET, BC_OUTLET = 1, 2, 3, 4
KIND_NONE, KIND_NO_REVERSE, KIND_DAMP_REVERSE, KIND_VELOCITY = -1, 1, 2, 3
KIND_MDOT, KIND_SLIP, KIND_STAGNATION, KIND_FORCE_SUPERSONIC = 4, 5, 6, 7
THERMAL_FIXED_T, THERMAL_FIXED_Q, THERMAL_ADIABATIC = 1, 2, 3
SPEC_NONE, SPEC_RHO, SPEC_P, SPEC_T, SPEC_STATE = 0, 1, 2, 3, 4
LIMITER_NONE, LIMITER_VK, LIMITER_BJ, LIMITER_MINMOD = -1, 1, 2, 3
ORDER_FIRST, ORDER_SECOND = 1, 2
FLUX_ROE, FLUX_VAN_LEER, FLUX_AUSM_PLUS_UP, FLUX_SD_SLAU, FLUX_SW = 1, 2, 3, 4, 5
TURB_NONE, TURB_SA, TURB_KW = 0, 1, 2
RANS_KEPSILON, RANS_KOMEGA, RANS_BSL, RANS_SST = 1, 2, 3, 4
GRAD_CURVILINEAR, GRAD_LSQR, GRAD_GREENGAUSS = 1, 2, 3
INTERP_WTLI, INTERP_IDW, INTERP_SIMPLE = 1, 2, 3
VISC_CONSTANT, VISC_SUTHERLAND = 1, 2
COND_CONSTANT, COND_PRANDTL = 1, 2
DT_FIXED, DT_CFL_MAX, DT_CFL_LOCAL = 1, 2, 3
MAX_BC = 32

UNIV_GAS_CONST = 8314.47215  # src/material/material.h:26

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("xyz", _dp), ("n_cells", C.c_int), ("cell_ptr", _ip),
                ("cell_nodes", _ip), ("n_bc", C.c_int), ("bc_ptr", _ip), ("bc_nodes", _ip),
                ("dimension", C.c_int)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int), ("kind", C.c_int), ("specified", C.c_int), ("thermal", C.c_int),
                ("p_fixed", C.c_int), ("T_fixed", C.c_int), ("rho_fixed", C.c_int), ("V_fixed", C.c_int),
                ("qdot_fixed", C.c_int),
                ("p", C.c_double), ("T", C.c_double), ("rho", C.c_double), ("V", C.c_double * 3),
                ("p_total", C.c_double), ("T_total", C.c_double), ("mdot", C.c_double), ("qdot", C.c_double),
                ("include_in_loads", C.c_int), ("nut", C.c_double),
                ("turb_intensity", C.c_double), ("eddy_visc_ratio", C.c_double)]


class Options(C.Structure):
    _fields_ = [("gamma", C.c_double), ("Mw", C.c_double), ("Pref", C.c_double), ("Tref", C.c_double),
                ("visc_model", C.c_int), ("mu", C.c_double), ("sut_mu_ref", C.c_double),
                ("sut_T_ref", C.c_double), ("sut_S", C.c_double),
                ("cond_model", C.c_int), ("lambda_", C.c_double), ("Pr", C.c_double),
                ("order", C.c_int), ("jac_order", C.c_int), ("limiter", C.c_int), ("convective_flux", C.c_int),
                ("limiter_threshold", C.c_double), ("Minf", C.c_double), ("wdiss", C.c_double),
                ("bl_height", C.c_double),
                ("grad_hex", C.c_int), ("grad_prism", C.c_int), ("grad_other", C.c_int),
                ("interp_method", C.c_int), ("stencil_size", C.c_int), ("skewness_tol", C.c_double),
                ("moment_center", C.c_double * 3),
                ("n_bc", C.c_int), ("bc", BC * MAX_BC), ("turbulence", C.c_int), ("Pr_t", C.c_double),
                ("rans_model", C.c_int), ("k_low", C.c_double), ("k_high", C.c_double), ("omega_low", C.c_double),
                ("visc_ratio_limit", C.c_double)]


class Info(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_cells", C.c_int), ("n_ghosts", C.c_int), ("n_faces", C.c_int),
                ("n_bfaces", C.c_int), ("ld_cell", C.c_int), ("ld_face", C.c_int), ("n_gradmap", C.c_int),
                ("n_faceavg", C.c_int), ("n_nbr", C.c_int), ("n_halffaces", C.c_int), ("nv", C.c_int),
                ("n_nodeavg", C.c_int)]


JAC_FD, JAC_JVP_FROZEN, JAC_JVP_FULL = 1, 2, 3


class NewtonOpts(C.Structure):
    _fields_ = [("dt_type", C.c_int), ("cfl", C.c_double), ("jacobian", C.c_int), ("krylov_fp32", C.c_int),
                ("restart", C.c_int), ("max_iters", C.c_int), ("rtol", C.c_double), ("abstol", C.c_double),
                ("knoll_guess", C.c_int), ("precond", C.c_int), ("sweeps", C.c_int), ("max_rel_update", C.c_double)]


class NewtonStats(C.Structure):
    _fields_ = [("res_norm", C.c_double), ("update_norms", C.c_double * 3), ("krylov_iters", C.c_int),
                ("krylov_converged", C.c_int), ("krylov_rnorm", C.c_double), ("krylov_r0norm", C.c_double),
                ("ms_residual", C.c_double), ("ms_jacobian", C.c_double), ("ms_krylov", C.c_double),
                ("ms_update", C.c_double), ("nut_norm", C.c_double), ("relax", C.c_double),
                ("rans_iters", C.c_int), ("rans_norms", C.c_double * 2)]


# the BC values of fcfd_vjp_params, in its order
BC_PARAMS = ("p", "T", "rho", "Vx", "Vy", "Vz", "p_total", "T_total", "mdot", "qdot", "nut")


class RansOpts(C.Structure):
    _fields_ = [("restart", C.c_int), ("max_iters", C.c_int), ("rtol", C.c_double), ("abstol", C.c_double)]


class RansStats(C.Structure):
    _fields_ = [("krylov_iters", C.c_int), ("krylov_converged", C.c_int), ("krylov_rnorm", C.c_double),
                ("krylov_r0norm", C.c_double), ("norms", C.c_double * 2)]


def rans_opts(restart=30, max_iters=10, rtol=1e-6, abstol=1e-12):
    """Defaults follow the reference's turbulence section (read_inputs.cc:113-115)."""
    return RansOpts(restart=restart, max_iters=max_iters, rtol=rtol, abstol=abstol)


PC_BLOCK_JACOBI, PC_BLOCK_SGS, PC_MULTIGRID = 0, 1, 2
PSEUDO_PC_NONE, PSEUDO_PC_WS95 = 0, 1


class PseudoOpts(C.Structure):
    _fields_ = [("dtau_type", C.c_int), ("dtau", C.c_double), ("preconditioner", C.c_int), ("ps_step", C.c_int),
                ("Wn_dev", C.c_void_p)]


def newton_opts(cfl=10.0, jacobian=JAC_FD, fp32=False, restart=30, max_iters=10, rtol=1e-6, abstol=1e-12, knoll=True,
                precond=PC_BLOCK_JACOBI, sweeps=1, max_rel_update=0.0):
    """Defaults follow the reference (read_inputs.cc:125-127, ns_petsc_functions.cc:74-79)."""
    return NewtonOpts(dt_type=DT_CFL_LOCAL, cfl=cfl, jacobian=jacobian, krylov_fp32=int(fp32), restart=restart,
                      max_iters=max_iters, rtol=rtol, abstol=abstol, knoll_guess=int(knoll), precond=precond, sweeps=sweeps,
                      max_rel_update=max_rel_update)


class MetricsOut(C.Structure):
    _fields_ = [("face_parent", _ip), ("face_neighbor", _ip), ("face_bc", _ip),
                ("face_node_ptr", _ip), ("face_nodes", _ip),
                ("face_centroid", _dp), ("face_normal", _dp), ("face_area", _dp), ("face_tangent1", _dp),
                ("cell_centroid", _dp), ("cell_volume", _dp), ("cell_lengthscale", _dp), ("wall_distance", _dp),
                ("cell_face_ptr", _ip), ("cell_faces", _ip), ("nbr_ptr", _ip), ("nbr", _ip),
                ("grad_ptr", _ip), ("grad_idx", _ip), ("grad_w", _dp),
                ("avg_ptr", _ip), ("avg_idx", _ip), ("avg_w", _dp),
                ("cell_perm", _ip), ("grid_lengthscale", _dp),
                ("node_xyz", _dp), ("node_avg_ptr", _ip), ("node_avg_idx", _ip), ("node_avg_w", _dp)]


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def make_mesh(raw):
    """Build a Mesh struct from a raw-mesh dict; keeps the arrays alive on the struct."""
    xyz = np.ascontiguousarray(raw["xyz"], dtype=np.float64).reshape(-1)
    cp = np.ascontiguousarray(raw["cell_ptr"], dtype=np.int32)
    cn = np.ascontiguousarray(raw["cell_nodes"], dtype=np.int32)
    bp = np.ascontiguousarray(raw["bc_ptr"], dtype=np.int32)
    bn = np.ascontiguousarray(raw["bc_nodes"], dtype=np.int32)
    m = Mesh(n_nodes=len(xyz) // 3, xyz=dptr(xyz), n_cells=len(cp) - 1, cell_ptr=iptr(cp),
             cell_nodes=iptr(cn), n_bc=len(bp) - 1, bc_ptr=iptr(bp), bc_nodes=iptr(bn),
             dimension=int(raw["dimension"]))
    m._keep = (xyz, cp, cn, bp, bn)
    return m


def alloc_metrics(info, full=True):
    """Allocate numpy buffers for every MetricsOut field; returns (struct, dict of arrays)."""
    F, N, B = info.n_faces, info.n_cells, info.n_ghosts
    T = N + B
    arrs = dict(
        face_parent=np.zeros(F, np.int32), face_neighbor=np.zeros(F, np.int32), face_bc=np.zeros(F, np.int32),
        face_node_ptr=np.zeros(F + 1, np.int32), face_nodes=np.zeros(4 * F, np.int32),
        face_centroid=np.zeros(3 * F), face_normal=np.zeros(3 * F), face_area=np.zeros(F),
        face_tangent1=np.zeros(3 * F),
        cell_centroid=np.zeros(3 * T), cell_volume=np.zeros(T), cell_lengthscale=np.zeros(N),
        wall_distance=np.zeros(T),
        cell_face_ptr=np.zeros(N + 1, np.int32), cell_faces=np.zeros(info.n_halffaces, np.int32),
        nbr_ptr=np.zeros(T + 1, np.int32), nbr=np.zeros(info.n_nbr, np.int32),
        grad_ptr=np.zeros(N + 1, np.int32), grad_idx=np.zeros(info.n_gradmap, np.int32),
        grad_w=np.zeros(3 * info.n_gradmap),
        avg_ptr=np.zeros(F + 1, np.int32), avg_idx=np.zeros(info.n_faceavg, np.int32),
        avg_w=np.zeros(info.n_faceavg),
        cell_perm=np.zeros(T, np.int32), grid_lengthscale=np.zeros(1),
        node_xyz=np.zeros(3 * info.n_nodes), node_avg_ptr=np.zeros(info.n_nodes + 1, np.int32),
        node_avg_idx=np.zeros(max(info.n_nodeavg, 1), np.int32), node_avg_w=np.zeros(max(info.n_nodeavg, 1)))
    m = MetricsOut()
    for k, a in arrs.items():
        setattr(m, k, iptr(a) if a.dtype == np.int32 else dptr(a))
    return m, arrs
