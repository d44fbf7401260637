/*
 * fcfd_gpu.h -- C-ABI of the MI355X (gfx950) finite-volume residual path.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes.
 * Every entry point returns FCFD_OK (0) on success or a negative FCFD_E*
 * code; the text of the last error is available from fcfd_last_error().
 * Nothing here calls exit() or aborts (the reference exits on input errors,
 * src/navier_stokes/ns_apply_bcs.cc:55-58, and aborts on divergence,
 * src/navier_stokes/ns.cc:350-355; both become return codes).
 *
 * Reference interface replaced (Free CFD 2.0, jackxiang/freecfd):
 *   mesh set-up     Grid::setup (src/grid/grid.cc:55), set_bcs (src/set_bcs.cc:35),
 *                   set_lengthScales (src/main.cc:456), face_interpolation_weights
 *                   (src/face_interpolation_weights.cc:32), gradient_maps
 *                   (src/gradient_maps.cc:39)                     -> fcfd_create
 *   per step        NavierStokes::update_boundaries (declared src/navier_stokes/ns.h:165),
 *                   calc_cell_grads (ns.cc:301), calc_limiter (ns_limiters.cc:26),
 *                   assemble_linear_system residual part (ns_assemble_linear_system.cc:75-140)
 *                                                                  -> fcfd_residual
 *                   get_jacobians + MatSetValues (ns_assemble_linear_system.cc:144-267),
 *                   time_terms (ns_time_terms.cc:107)              -> fcfd_fd_jacobian
 *                   update_time_step (src/time_step.cc:143)       -> fcfd_time_step
 *                   petsc_solve + update_variables (ns_petsc_functions.cc:85, ns.cc:338)
 *                                                                  -> fcfd_newton_step
 *                   surface loads of the face loop (ns_assemble_linear_system.cc:109-119)
 *                                                                  -> fcfd_loads
 *   build-only      tangent and adjoint of the residual           -> fcfd_jvp, fcfd_vjp
 *                   steady adjoint solve, BC parameter updates    -> fcfd_adjoint_step, fcfd_set_bc
 *
 * Numbering.  Host arrays passed to or returned by fcfd_* in "reference
 * numbering" use the cell, face and node order the reference's Grid::setup
 * produces on one rank (cells in input order, faces in creation order of
 * src/grid/grid_create_elements.cc:184-442, boundary ghosts appended per BC
 * region in face order, grid_create_elements.cc:702-743).  Device arrays use
 * the context's internal numbering (cells along a space-filling curve, or in
 * reverse Cuthill-McKee order with the environment variable
 * FCFD_CELL_ORDER=rcm; results do not depend on it) and a
 * structure-of-arrays layout: variable v of internal cell i lives at
 * ptr[v*ld + i], where ld = fcfd_info.ld_cell (a multiple of 64).
 */
#ifndef FCFD_GPU_H
#define FCFD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FCFD_OK            0
#define FCFD_EINVAL       -1   /* bad argument / inconsistent input            */
#define FCFD_EMESH        -2   /* mesh cannot be processed (unsupported cell)   */
#define FCFD_EBC          -3   /* boundary-condition input error                */
#define FCFD_EHIP         -4   /* HIP runtime error                              */
#define FCFD_EDIVERGED    -5   /* NaN/Inf detected in an update (ns.cc:350-355) */
#define FCFD_ENOTSUP      -6   /* option not supported by this build             */
#define FCFD_ENOMEM       -7

/* ---- option values: same numbers as the reference headers ---- */
/* BC types/kinds/specification (src/bc.h:28-51) */
#define FCFD_BC_SYMMETRY 1
#define FCFD_BC_WALL     2
#define FCFD_BC_INLET    3
#define FCFD_BC_OUTLET   4
#define FCFD_KIND_NONE        -1
#define FCFD_KIND_NO_REVERSE   1
#define FCFD_KIND_DAMP_REVERSE 2
#define FCFD_KIND_VELOCITY     3
#define FCFD_KIND_MDOT         4
#define FCFD_KIND_SLIP         5
#define FCFD_KIND_STAGNATION   6
#define FCFD_KIND_FORCE_SUPERSONIC 7
#define FCFD_THERMAL_FIXED_T   1
#define FCFD_THERMAL_FIXED_Q   2
#define FCFD_THERMAL_ADIABATIC 3
#define FCFD_SPEC_NONE  0
#define FCFD_SPEC_RHO   1
#define FCFD_SPEC_P     2
#define FCFD_SPEC_T     3
#define FCFD_SPEC_STATE 4
/* numerics (src/navier_stokes/ns.h:40-55) */
#define FCFD_LIMITER_NONE   -1
#define FCFD_LIMITER_VK      1
#define FCFD_LIMITER_BJ      2
#define FCFD_LIMITER_MINMOD  3
#define FCFD_ORDER_FIRST     1
#define FCFD_ORDER_SECOND    2
#define FCFD_FLUX_ROE        1
#define FCFD_FLUX_VAN_LEER   2
#define FCFD_FLUX_AUSM_PLUS_UP 3
#define FCFD_FLUX_SD_SLAU    4
#define FCFD_FLUX_SW         5
/* gradient methods (src/gradient_maps.cc:26-28) */
#define FCFD_GRAD_CURVILINEAR 1
#define FCFD_GRAD_LSQR        2
#define FCFD_GRAD_GREENGAUSS  3   /* face sum of face values (variable.h:193-216): residual,
                                     tangent, both Newton modes, the adjoint and the
                                     BC-value derivatives (a face value a BC holds
                                     fixed is that BC's value) */
/* interpolation (src/interpolate/interpolate.h).  A stencil above 2 (WTLI
 * fits, or IDW over the extended stencil) gives faces whose viscous face
 * gradients average more than their two cells: residual, FD Jacobian,
 * tangent, both Newton modes and the adjoint (fcfd_vjp transposes the
 * average; a ghost in a stencil passes its weight to its parent, whose
 * gradient it copies).  fcfd_create returns FCFD_ENOTSUP for such a viscous
 * context with the segregated two-equation model */
#define FCFD_INTERP_WTLI   1
#define FCFD_INTERP_IDW    2
#define FCFD_INTERP_SIMPLE 3
/* material models (src/material/material.h:28-35) */
#define FCFD_VISC_CONSTANT    1
#define FCFD_VISC_SUTHERLAND  2
#define FCFD_COND_CONSTANT    1
#define FCFD_COND_PRANDTL     2
/* time step type (src/time_step.h) */
#define FCFD_DT_FIXED     1
#define FCFD_DT_CFL_MAX   2
#define FCFD_DT_CFL_LOCAL 3

/* turbulence model */
#define FCFD_TURB_NONE 0
#define FCFD_TURB_SA   1   /* build-only: Spalart-Allmaras nu~ as a coupled sixth equation   */
#define FCFD_TURB_KW   2   /* the reference's segregated two-equation RANS (src/rans/): k and
                              omega solved after each NS step (fcfd_rans_step), mu_t frozen
                              in the NS face loop (ns_diffusive_face_flux.cc:36-39)          */
/* two-equation model (src/rans/rans.h:44-47; "k-epsilon", "k-omega", "sst"
 * in rans.cc:54-60; BSL has no input keyword there) */
#define FCFD_RANS_KEPSILON 1
#define FCFD_RANS_KOMEGA   2
#define FCFD_RANS_BSL      3
#define FCFD_RANS_SST      4

#define FCFD_MAX_BC 32

/* Raw unstructured mesh, the content of the reference's GridRawData for
 * cell-based input (src/grid/grid.h:61-78).  Cells use CGNS node order for
 * TETRA_4 / PYRA_5 / PENTA_6 / HEXA_8.  bc node lists are raw node ids; each
 * list is a set (order and duplicates do not matter). */
typedef struct fcfd_mesh {
    int n_nodes;
    const double* xyz;          /* 3*n_nodes, x,y,z interleaved               */
    int n_cells;
    const int* cell_ptr;        /* n_cells+1 offsets into cell_nodes          */
    const int* cell_nodes;      /* raw node ids                               */
    int n_bc;                   /* number of BC regions (node sets)           */
    const int* bc_ptr;          /* n_bc+1 offsets into bc_nodes               */
    const int* bc_nodes;        /* raw node ids                               */
    int dimension;              /* 2 or 3 (grid_N{dimension=...})             */
} fcfd_mesh;

/* One boundary region; fields mirror BCregion (src/bc.h:54-63) and the
 * per-BC values set in NavierStokes::set_bcs (ns_set_bcs.cc:25-150). */
typedef struct fcfd_bc {
    int type;         /* FCFD_BC_*                                  */
    int kind;         /* FCFD_KIND_*                                */
    int specified;    /* FCFD_SPEC_*                                */
    int thermal;      /* FCFD_THERMAL_*                             */
    int p_fixed, T_fixed, rho_fixed, V_fixed, qdot_fixed;  /* Variable::fixedonBC */
    double p, T, rho, V[3], p_total, T_total, mdot, qdot;
    int include_in_loads;
    double nut;       /* SA working variable nu~ at an inlet (m^2/s); build-only */
    /* two-equation RANS at an inlet: turbulenceintensity and eddyviscosityratio
     * (read_inputs.cc:110-111, defaults 1e-4 and 0.1; rans_set_bcs.cc:33-64) */
    double turb_intensity, eddy_visc_ratio;
} fcfd_bc;

typedef struct fcfd_options {
    /* material (src/material/material.cc:29-106) */
    double gamma, Mw, Pref, Tref;
    int visc_model;  double mu, sut_mu_ref, sut_T_ref, sut_S;
    int cond_model;  double lambda, Pr;
    /* navierstokes section (src/read_inputs.cc:125-134; ns.cc:33-86) */
    int order, jac_order, limiter, convective_flux;
    double limiter_threshold, Minf, wdiss, bl_height;
    /* gradients and interpolation sections (read_inputs.cc:47-54) */
    int grad_hex, grad_prism, grad_other;
    int interp_method, stencil_size;
    double skewness_tol;
    /* loads */
    double moment_center[3];
    int n_bc;
    fcfd_bc bc[FCFD_MAX_BC];
    /* turbulence (build-only; the reference's own RANS path is segregated
     * k-omega/SST): FCFD_TURB_SA adds nu~ as a sixth variable, coupled */
    int turbulence;
    double Pr_t;      /* turbulent Prandtl number (the reference's turbulentPr) */
    /* turbulence section of the two-equation RANS (read_inputs.cc:117-121):
     * model (FCFD_RANS_*, default SST), klowlimit (1e-10), khighlimit (1e8),
     * omegalowlimit (1e-2), viscosityratiolimit (1e8) */
    int rans_model;
    double k_low, k_high, omega_low, visc_ratio_limit;
} fcfd_options;

typedef struct fcfd_info {
    int n_nodes, n_cells, n_ghosts, n_faces, n_bfaces;
    int ld_cell;         /* padded leading dimension of cell SoA arrays (>= n_cells+n_ghosts) */
    int ld_face;
    int n_gradmap;       /* total gradient-map entries over interior cells */
    int n_faceavg;       /* total face-average entries over faces          */
    int n_nbr;           /* total neighborCells entries over all cells     */
    int n_halffaces;     /* sum over interior cells of faces per cell      */
    int nv;              /* equations per cell (5, or 6 with the SA model) */
    int n_nodeavg;       /* total node-average entries over nodes          */
} fcfd_info;

/* Metrics in reference numbering for bit-exact checks of the set-up.  Every
 * pointer may be NULL (skipped); sizes come from fcfd_info. */
typedef struct fcfd_metrics_out {
    int* face_parent; int* face_neighbor; int* face_bc;   /* n_faces          */
    int* face_node_ptr; int* face_nodes;                  /* n_faces+1, <=4*n_faces */
    double* face_centroid; double* face_normal; double* face_area; /* 3F,3F,F */
    double* face_tangent1;                                /* 3F (ns_assemble_linear_system.cc:337-341) */
    double* cell_centroid; double* cell_volume;           /* 3(N+B), N+B      */
    double* cell_lengthscale;                             /* N                */
    double* wall_distance;                                /* N+B              */
    int* cell_face_ptr; int* cell_faces;                  /* N+1, halffaces   */
    int* nbr_ptr; int* nbr;                               /* N+B+1, n_nbr     */
    int* grad_ptr; int* grad_idx; double* grad_w;         /* N+1, G, 3G       */
    int* avg_ptr; int* avg_idx; double* avg_w;            /* F+1, A, A        */
    int* cell_perm;                                       /* N+B: internal index of reference cell */
    double* grid_lengthscale;                             /* 1                */
    /* nodes in the reference's local order (first touch in cell order) and
     * their interpolation weights, node.average of
     * node_interpolation_weights (src/node_interpolation_weights.cc:32-118):
     * keys are cells in reference numbering (ghosts included), ascending */
    double* node_xyz;                                     /* 3 n_nodes        */
    int* node_avg_ptr; int* node_avg_idx; double* node_avg_w; /* n_nodes+1, n_nodeavg, n_nodeavg */
} fcfd_metrics_out;

typedef struct fcfd_ctx fcfd_ctx;

/* Build the context: host set-up (connectivity, geometry, ghosts, BCs,
 * length scales, interpolation weights, gradient maps, renumbering) and
 * device upload.  Replaces Grid::setup .. gradient_maps (main.cc:181-189). */
int  fcfd_create(const fcfd_mesh* mesh, const fcfd_options* opt, fcfd_ctx** out);
/* Host-only set-up (no GPU needed): same as fcfd_create without device
 * allocation; the context then only serves fcfd_info/fcfd_metrics. */
int  fcfd_create_host(const fcfd_mesh* mesh, const fcfd_options* opt, fcfd_ctx** out);
void fcfd_destroy(fcfd_ctx* ctx);
const char* fcfd_last_error(const fcfd_ctx* ctx);
int  fcfd_get_info(const fcfd_ctx* ctx, fcfd_info* info);
int  fcfd_metrics(const fcfd_ctx* ctx, fcfd_metrics_out* out);

/* Layout helpers.  W_host is N x nv in reference numbering, AoS
 * (p,u,v,w,T per cell, the order of Variable p, V, T in ns.h:104-106, then
 * nu~ with the SA model; nv = fcfd_info.nv).  W_dev is a device SoA array
 * [nv][ld_cell]. */
int  fcfd_state_to_device(fcfd_ctx* ctx, const double* W_host, double* W_dev);
int  fcfd_cells_to_host(fcfd_ctx* ctx, const double* X_dev, int nvar, int include_ghosts, double* X_host);

/* FP64 residual R(W): ghost update (update_boundaries, decided semantics),
 * gradients (calc_cell_grads), limiters (calc_limiter) and the face-flux
 * loop (ns_assemble_linear_system.cc:75-140).  W_dev is [nv][ld_cell]; its
 * interior slots [0,N) are read and its boundary-ghost slots [N,N+B) are
 * written with the ghost state (the reference keeps ghosts in the same
 * Variable arrays, variable.h:84-89).  R_dev is [nv][ld_cell], interior cells:
 * R_c = sum over the faces of c, in reference face order, of +-(D - C).
 * Stream-ordered on `stream` (a hipStream_t, or NULL for the null stream);
 * no host synchronisation. */
int  fcfd_residual(fcfd_ctx* ctx, double* W_dev, double* R_dev, void* stream);

/* Boundary ghosts, gradients and limiters of W_dev alone: replaces
 * NavierStokes::update_boundaries (declared src/navier_stokes/ns.h:165),
 * calc_cell_grads (ns.cc:301-336) and calc_limiter (ns_limiters.cc:26-248,
 * with find_min_max ns.cc:400-435 for VK).  Writes the ghost slots of W_dev,
 * the gradients (ghosts carry their parent's) into G_dev and the limiters of
 * interior cells into PHI_dev, both in the row-pair layout below and sized
 * 2*ceil(3nv/2)*ld_cell and 2*ceil(nv/2)*ld_cell doubles; a NULL G_dev or
 * PHI_dev selects the context's own buffer.  PHI_dev is not written when no
 * limiter runs (first order, or limiter NONE: phi is then the constant 0 or 1
 * of ns.cc:282-287).  Stream-ordered; no host synchronisation. */
int  fcfd_gradients_limiters(fcfd_ctx* ctx, double* W_dev, double* G_dev, double* PHI_dev, void* stream);

/* Gradients (3nv components per cell: dp/dx,dp/dy,dp/dz, du/.., dv/.., dw/..,
 * dT/.. [, dnu~/..]; ghosts carry their parent's) and limiters (nv per cell)
 * of the last fcfd_residual call, or fcfd_gradients_limiters call into the
 * context's own buffers; device pointers owned by the context.
 * Stored in row pairs: component k of cell i is at ((k/2)*ld_cell + i)*2 +
 * k%2 (fcfd_gradients_to_host returns them per cell in reference order). */
int  fcfd_internal_buffers(fcfd_ctx* ctx, const double** grad, const double** phi);

/* The same, unpacked to the host in reference numbering: G_host (N+B) x 3nv,
 * PHI_host N x nv (either may be NULL).  Synchronises the device. */
int  fcfd_gradients_to_host(fcfd_ctx* ctx, double* G_host, double* PHI_host);

/* Per-stage timing of fcfd_residual with HIP events recorded on the launch
 * stream around its four stages (ghosts, gradients+limiters, face fluxes,
 * ordered cell sums): stage_ms must hold 4 values.
 * Returns the accumulated milliseconds per stage and the number of timed
 * calls; enable = 1 starts (and resets), 0 stops (and resets), -1 only
 * reads.  No counterpart in the reference (its only timer is the wall clock
 * of main.cc:445-448). */
int  fcfd_profile(fcfd_ctx* ctx, int enable, double* stage_ms, long long* calls);

/* Synchronises and reports run-time BC errors flagged by kernels since the
 * last clear (velocity inlet turned supersonic without a complete state,
 * ns_apply_bcs.cc:54-59 -> FCFD_EBC). */
int  fcfd_status(fcfd_ctx* ctx, int clear);

/* The same time step as fcfd_time_step, from an entry point that allocates
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc), where fcfd_time_step's first call
 * allocates its reduction buffer.  Its scratch (the CFL_MAX minimum, the
 * largest CFL number and 256 block partials: 258 doubles, whatever the mesh
 * size, so no input is too large for it) is one static device array per GPU,
 * shared by every call on that GPU: calls that overlap on one GPU race.  The
 * largest CFL number goes to the device pointer max_cfl_dev (may be NULL),
 * stream-ordered, with no host synchronisation; dt_dev is written as by
 * fcfd_time_step.  Calls no hipMalloc/hipFree and no allocating library
 * routine. */
int  fcfd_time_step_noalloc(fcfd_ctx* ctx, const double* W_dev, int dt_type, double cfl, double* dt_dev, double* max_cfl_dev,
                            void* stream);

/* Reference-equivalent FP64 finite-difference Jacobian of the face loop
 * (get_jacobians + MatSetValues, ns_assemble_linear_system.cc:144-267) and,
 * when dt_dev is not NULL, the time terms P*V/dt (ns_time_terms.cc:147-163).
 * Recomputes ghosts, gradients and limiters of W_dev first.  dWprev_dev
 * ([nv][ld_cell], the previous Newton update that sets the FD step,
 * :217-219) may be NULL (zero update).  The result stays in the context:
 * per face JL and JR (nv x nv entries per face: entry k = nv*var+eq of face f
 * at ((k/2)*F + f)*2 + k%2, in internal face order, entries stored in row
 * pairs) and diagonal blocks D (row-major nv x nv, entry k of cell i at
 * ((k/2)*ld_cell + i)*2 + k%2); fcfd_jacobian_to_host returns them per block
 * in reference order.  The operator row P, column N block is -JR^T of the
 * face (P,N); row N, column P is +JL^T. */
int  fcfd_fd_jacobian(fcfd_ctx* ctx, double* W_dev, const double* dWprev_dev, const double* dt_dev, void* stream);
int  fcfd_jacobian_buffers(fcfd_ctx* ctx, const double** JL, const double** JR, const double** D);
/* Copies the last Jacobian to the host in reference numbering: JL, JR as
 * [n_faces][5 var][5 eq], D as [n_cells][5 row][5 col].  Any may be NULL. */
int  fcfd_jacobian_to_host(fcfd_ctx* ctx, double* JL_host, double* JR_host, double* D_host);

/* ---- implicit Newton-Krylov step ---- */
#define FCFD_JAC_FD          1   /* reference FD Jacobian, assembled per face        */
#define FCFD_JAC_JVP_FROZEN  2   /* matrix-free exact tangent, gradients/limiters frozen */
#define FCFD_JAC_JVP_FULL    3   /* matrix-free exact tangent of the full residual    */

typedef struct fcfd_newton_opts {
    int dt_type;          /* FCFD_DT_CFL_LOCAL, FCFD_DT_CFL_MAX or FCFD_DT_FIXED
                             (update_time_step, time_step.cc:155-200) */
    double cfl;           /* the CFL number; the step size for FCFD_DT_FIXED */
    int jacobian;         /* FCFD_JAC_FD: assembled reference FD Jacobian; FCFD_JAC_JVP_*:
                             matrix-free tangent operator (P V/dt) x - J x.  Every
                             preconditioner is built from the blocks of the FD Jacobian +
                             time terms (see precond).                                 */
    int krylov_fp32;      /* 0: FP64 GMRES basis, FP64 preconditioner.  1: FP32 GMRES basis
                             and Gram-Schmidt; block Jacobi and block SGS in FP32; the
                             multigrid V-cycle stays FP64 (vectors converted at its
                             boundary); with FCFD_JAC_FD the BSR matvec runs in FP32 on
                             FP32 copies of the blocks, with FCFD_JAC_JVP_* the tangent
                             runs in FP64 on the converted vector                      */
    int restart;          /* GMRES restart length (the reference uses FGMRES(300))       */
    int max_iters;        /* maximumiterations (default 10, read_inputs.cc:127)          */
    double rtol, abstol;  /* relativetolerance / absolutetolerance (read_inputs.cc:125-126) */
    int knoll_guess;      /* initial guess = M^{-1} b (KSPSetInitialGuessKnoll, ns_petsc_functions.cc:75) */
    int precond;          /* FCFD_PC_BLOCK_JACOBI (default); FCFD_PC_BLOCK_SGS: multicolour block
                             symmetric Gauss-Seidel with the FD Jacobian blocks; FCFD_PC_MULTIGRID:
                             one FP64 V-cycle of aggregation multigrid on the same blocks (SGS
                             smoothing, Galerkin coarse levels).  Build-only; the reference leaves
                             the PC to PETSc options                                         */
    int sweeps;           /* SGS sweeps per application / per level and V-cycle side (>= 1)  */
    double max_rel_update;/* > 0: scale the whole update so that no cell's absolute p or T
                             changes by more than this fraction (a step-length limit for
                             high CFL; build-only); 0: the reference's full update          */
} fcfd_newton_opts;
#define FCFD_PC_BLOCK_JACOBI 0
#define FCFD_PC_BLOCK_SGS    1
#define FCFD_PC_MULTIGRID    2

typedef struct fcfd_newton_stats {
    double res_norm;          /* ||R(W)||_2 over interior cells and the nv equations, before the update */
    double update_norms[3];   /* sqrt(sum dp^2/dt^2), sqrt(sum |dV|^2/dt^2), sqrt(sum dT^2/dt^2) (ns.cc:377-394) */
    int krylov_iters, krylov_converged;
    double krylov_rnorm, krylov_r0norm;
    double ms_residual, ms_jacobian, ms_krylov, ms_update;
    double nut_norm;          /* sqrt(sum dnu~^2/dt^2) with the SA model, else 0 */
    double relax;             /* factor applied to the update (1 unless max_rel_update limited it) */
    /* the segregated two-equation model's solve inside the step (FCFD_TURB_KW
     * after fcfd_rans_init / fcfd_rans_set; else 0): its Krylov iterations
     * and sqrt(sum dk^2/dt^2), sqrt(sum domega^2/dt^2) (rans.cc:206-222) */
    int rans_iters;
    double rans_norms[2];
} fcfd_newton_stats;

/* ---- segregated two-equation RANS (opt.turbulence = FCFD_TURB_KW) ---- */
typedef struct fcfd_rans_opts {
    int restart;          /* GMRES restart length (the reference uses FGMRES(300))                    */
    int max_iters;        /* turbulence maximumiterations (default 10, read_inputs.cc:115)         */
    double rtol, abstol;  /* turbulence relativetolerance / absolutetolerance (1e-6 / 1e-12,
                             read_inputs.cc:113-114)                                               */
} fcfd_rans_opts;
typedef struct fcfd_rans_stats {
    int krylov_iters, krylov_converged;
    double krylov_rnorm, krylov_r0norm;
    double norms[2];      /* sqrt(sum dk^2/dt^2), sqrt(sum domega^2/dt^2) of the limited update   */
} fcfd_rans_stats;

/* ---- dual (pseudo) time stepping ---- */
#define FCFD_PSEUDO_PC_NONE 0   /* pseudo-time block = cons2prim (ns_time_terms.cc:26-56) */
#define FCFD_PSEUDO_PC_WS95 1   /* Weiss-Smith low-Mach preconditioner (ns_time_terms.cc:58-105) */

typedef struct fcfd_pseudo_opts {
    int dtau_type;          /* FCFD_DT_FIXED, FCFD_DT_CFL_MAX or FCFD_DT_CFL_LOCAL for the pseudo step
                               (update_pseudo_time_step, src/time_step.cc:226-290)             */
    double dtau;            /* its CFL number, or the step size for FCFD_DT_FIXED             */
    int preconditioner;     /* FCFD_PSEUDO_PC_NONE or FCFD_PSEUDO_PC_WS95 (pseudotime.preconditioner) */
    int ps_step;            /* pseudo step within the physical step, from 1                   */
    const double* Wn_dev;   /* [nv][ld_cell] state at the physical time level n; read for ps_step > 1 */
} fcfd_pseudo_opts;

/* One pseudo step of dual time stepping (NavierStokes::solve inside the
 * pseudo-time loop of main.cc:322-345, with time_terms for
 * pseudotime.numberofsteps > 1, ns_time_terms.cc:107-186): as
 * fcfd_newton_step, with the physical step from opts (dt_type, cfl) and the
 * implicit operator P V/dt + P_ps V/dtau - dR/dW (P_ps: cons2prim or the
 * WS95 matrix) and, for ps_step > 1, the right-hand side R - (P V/dt)(W - Wn).
 * stats->res_norm is the norm of that right-hand side. */
int  fcfd_pseudo_step(fcfd_ctx* ctx, double* W_dev, const fcfd_newton_opts* opts, const fcfd_pseudo_opts* pseudo,
                      fcfd_newton_stats* stats, void* stream);

/* Tangent (Jacobian-vector product) of the residual: dR = (dR/dW) dW, exact
 * derivatives by forward-mode dual numbers through the residual kernels.
 * mode FCFD_JAC_JVP_FULL differentiates ghosts, gradients and limiters (the
 * Jacobian of the second-order residual; the ghost slots [N,N+B) of W_dev
 * and dW_dev are written); FCFD_JAC_JVP_FROZEN holds gradients, limiters and
 * ghost centres fixed -- the linearisation the reference FD Jacobian
 * approximates.  Arrays [nv][ld_cell].  No counterpart in the reference,
 * which only builds the FD Jacobian. */
int  fcfd_jvp(fcfd_ctx* ctx, double* W_dev, double* dW_dev, double* dR_dev, int mode, void* stream);

/* One implicit step (NavierStokes::solve without the gradient refresh,
 * ns.cc:104-118): residual, local time step, Jacobian (or matrix-free
 * tangent), time terms, GMRES with block-Jacobi, W += dW with the NaN/Inf
 * check (FCFD_EDIVERGED, ns.cc:350-355).  The update is kept in the context
 * as the previous update that sets the FD step of the next call.  W_dev is
 * [nv][ld_cell], updated in place; it is left unchanged when the update holds a
 * NaN/Inf (FCFD_EDIVERGED). */
int  fcfd_newton_step(fcfd_ctx* ctx, double* W_dev, const fcfd_newton_opts* opts, fcfd_newton_stats* stats, void* stream);

/* Time step of every cell (update_time_step, src/time_step.cc:143-224):
 * FCFD_DT_CFL_LOCAL (cfl L / (|V| + a) per cell), FCFD_DT_CFL_MAX (the
 * smallest of those for every cell) or FCFD_DT_FIXED (`cfl` is the step
 * size).  dt_dev is [ld_cell].  *max_cfl (may be NULL) as the reference: the
 * CFL number for CFL_LOCAL, else the largest (|V| + a) dt / L, which waits
 * for the stream. */
int  fcfd_time_step(fcfd_ctx* ctx, const double* W_dev, int dt_type, double cfl, double* dt_dev, double* max_cfl, void* stream);

/* Adjoint (vector-Jacobian product) of the residual, the exact transpose of
 * fcfd_jvp's FULL mode:
 *   Wbar = (dR/dW)^T Rbar + d/dW sum_{boundary faces f} fx_weight[bc(f)] . (D - C)_f
 * where D - C is the face's contribution to its parent cell's residual.  The
 * second term gives the state adjoint of a surface-load functional: the
 * force on BC b is -sum_f (D - C)_f[1..3], so a drag direction e on a wall
 * BC is fx_weight[b] = (0, -e_x, -e_y, -e_z, 0).  Rbar_dev and Wbar_dev are
 * [nv][ld_cell] device arrays (only interior cells of Rbar are read);
 * fx_weight_dev is a device [n_bc][5] array or NULL.  The ghost slots of
 * W_dev are written.  Deterministic (no atomics).  No counterpart in the
 * reference, which has no adjoint. */
int  fcfd_vjp(fcfd_ctx* ctx, double* W_dev, const double* Rbar_dev, const double* fx_weight_dev, double* Wbar_dev, void* stream);

/* The exact derivatives with respect to BC values (no counterpart in the
 * reference: its BC values are fixed inputs, src/set_bcs.cc:35,
 * ns_set_bcs.cc:25-180).  A BC's values enter the residual through its
 * ghosts' centre states (then the gradients and limiters of the cells next
 * to it), the right states of its faces (apply_bcs), a FIXED_Q wall's heat
 * flux and the SA inlet nu~; these derivatives follow every path, by the
 * tangent kernels with the BC's values as dual numbers (with Green-Gauss
 * gradients also the face values the BC holds fixed).
 *
 * fcfd_bc_tangent: dR_dev [nv][ld_cell] = dR/d(values of BC `index`) . dbc at
 * the fixed state W_dev, and (dloads_host may be NULL) the loads' derivative
 * in the layout of fcfd_loads (6 per BC).  dbc holds the direction in its
 * fields p, T, rho, V, p_total, T_total, mdot, qdot, nut (the others are
 * ignored).  The ghost slots of W_dev are written. */
int  fcfd_bc_tangent(fcfd_ctx* ctx, double* W_dev, int index, const fcfd_bc* dbc, double* dR_dev, double* dloads_host, void* stream);
/* The parameter part of the VJP (pbar): for every BC b and each of its
 * FCFD_BC_NPARAMS values (p, T, rho, V[0], V[1], V[2], p_total, T_total,
 * mdot, qdot, nut, in this order), pbar_host[FCFD_BC_NPARAMS b + q] =
 * <Rbar, dR/dv> + the derivative of sum over boundary faces of
 * fx_weight[bc][0..4] . (D - C) (fcfd_vjp's output term; fx_weight_dev may be
 * NULL).  With Rbar = the adjoint lam and fx_weight an output's load weight
 * this is lam . dR/dv + dJ/dv|_W, the adjoint sensitivity to each BC value.
 * Exact; forward mode, one tangent evaluation per BC value. */
#define FCFD_BC_NPARAMS 11
int  fcfd_vjp_params(fcfd_ctx* ctx, double* W_dev, const double* Rbar_dev, const double* fx_weight_dev, double* pbar_host,
                     void* stream);

/* Surface loads of the state W_dev (ns_assemble_linear_system.cc:109-119,
 * loads.h:33-39): for every BC with include_in_loads, force = sum over its
 * faces of (C - D)[1..3] and moment = sum (x_f - moment_center) x force_f,
 * summed in reference face order; BCs not included report zeros.  Writes
 * 6 * n_bc doubles (Fx Fy Fz Mx My Mz per BC) to loads_host after
 * synchronising the stream.  The ghost slots of W_dev are written. */
int  fcfd_loads(fcfd_ctx* ctx, double* W_dev, double* loads_host, void* stream);

/* Surface quantities of the face loop at the state W_dev
 * (ns_assemble_linear_system.cc:122-127), in reference face order on the
 * host: mdot_host[F] = (C - D)[0] / A on every face (mdot.face); on boundary
 * faces qdot_host[F] = (C - D)[4] / A, or the given qdot where qdot_fixed
 * (qdot.bc), and tau_host[3F] = -D[1..3] / A (tau.bc); both 0 on interior
 * faces.  The surface writer's mdot, qdot and tau (write_surface_output.cc:
 * 266-290, 390).  Any output may be NULL.  Recomputes ghosts, gradients and
 * limiters of W_dev (its ghost slots are written); synchronises the stream. */
int  fcfd_surface(fcfd_ctx* ctx, double* W_dev, double* mdot_host, double* qdot_host, double* tau_host, void* stream);

/* The convective flux of a batch of n faces from given states: the leaf
 * functions roe_flux, AUSMplusUP_flux, vanLeer_flux, SD_SLAU_flux and
 * Stegger_Warming_flux with convective_face_flux's rotation to xyz and
 * area (ns_convective_face_flux.cc:25-59; ns_roe.cc:25-98,
 * ns_ausm_plus_up.cc:35-152, ns_vanleer.cc:25-66, ns_sd_slau.cc:37-94,
 * ns_stegger_warming.cc:23-145).  Device arrays, SoA: WL_dev/WR_dev [5][n]
 * primitive p, u, v, w, T of the left and right states; normal_dev and
 * tangent1_dev [3][n] (unit vectors; t2 = norm(n x t1)); area_dev [n];
 * diss_dev and wall_dist_dev [n] enter Steger-Warming only and may be NULL
 * (then 1 and 1e20).  Each state is finished as the face loop does it (rho
 * by the EOS, a, H, Vn in the face frame) with the context's material.
 * Writes flux_dev [5][n] (area-weighted, xyz) and, if not NULL,
 * weightL_dev [n] (weightL.face as the flux function leaves it).  flux is
 * an FCFD_FLUX_* value, independent of the context's own scheme.  Stream
 * ordered; n = 0 does nothing. */
int  fcfd_convective_flux(fcfd_ctx* ctx, int flux, int n, const double* WL_dev, const double* WR_dev, const double* normal_dev,
                          const double* tangent1_dev, const double* area_dev, const double* diss_dev,
                          const double* wall_dist_dev, double* flux_dev, double* weightL_dev, void* stream);

/* One pseudo-time step of the steady adjoint (dR/dW)^T lam = -g at the state
 * W_dev (g = dJ/dW of an output J, e.g. from fcfd_vjp with a load weight and
 * Rbar = 0): with rho = (dR/dW)^T lam + g, solves ((P V/dt) - dR/dW)^T delta
 * = rho by FP64 GMRES with the exact transpose (fcfd_vjp) as the operator and,
 * as the preconditioner, the transposed FD block Jacobi (FCFD_PC_BLOCK_JACOBI)
 * `sweeps` symmetric multicolour block Gauss-Seidel sweeps on the
 * transposed FD operator (FCFD_PC_BLOCK_SGS), or one V-cycle of the
 * aggregation multigrid on the transposed FD operator (FCFD_PC_MULTIGRID:
 * every level's blocks transposed) of the same step, then lam += delta.  As CFL grows it tends to the Newton solve of the adjoint
 * system, which is linear: lam converges at the rate of the Krylov solves.
 * opts: cfl, restart, max_iters, rtol, abstol, knoll_guess, precond, sweeps;
 * krylov_fp32 must be 0.  stats->res_norm = ||rho||_2 before
 * the step.  g_dev and lam_dev are [nv][ld_cell] (lam updated in place);
 * the ghost slots of W_dev are written.  No counterpart in the reference. */
int  fcfd_adjoint_step(fcfd_ctx* ctx, double* W_dev, const double* g_dev, double* lam_dev, const fcfd_newton_opts* opts,
                       fcfd_newton_stats* stats, void* stream);

/* Replace the values of BC `index` (p, T, rho, V, p_total, T_total, mdot,
 * qdot, nut) in a live context, the mesh and its set-up unchanged: type,
 * kind, specified and thermal must equal the ones given at create
 * (FCFD_EBC otherwise), and the change may not switch the diffusive flux on
 * or off (a FIXED_Q wall's qdot between 0 and non-zero on an otherwise
 * inviscid case: FCFD_EBC).  Synchronises the device first, so no work of the
 * context may be in flight on another thread.  For parameter studies and
 * the parameter derivatives of an adjoint; the reference reads BC values
 * once from its input file (src/set_bcs.cc:35, ns_set_bcs.cc:25-180). */
int  fcfd_set_bc(fcfd_ctx* ctx, int index, const fcfd_bc* bc);

/* ---- segregated two-equation RANS (src/rans/), FCFD_TURB_KW ----
 * The context holds k, omega, mu_t and their gradients on the device (the
 * RANS object of the reference); the NS face loop of every entry point takes
 * its mu_t from them (rans.mu_t.face, ns_diffusive_face_flux.cc:36-39), frozen
 * in the NS Jacobian and tangents as in the reference.  Until fcfd_rans_init
 * or fcfd_rans_set runs, mu_t is 0.  Every call below returns FCFD_EINVAL on
 * a context without the model. */

/* RANS::initialize (rans.cc:70-81): apply_initial_conditions over the whole
 * domain with the given turbulence intensity and eddy viscosity ratio
 * (rans.cc:120-170), set_bcs (the inlets' k, omega, mu_t from the NS face
 * values of each inlet's first face, rans_set_bcs.cc:39-79), ghosts, eddy
 * viscosity (rans_update_eddy_viscosity.cc:25-59) and gradients
 * (rans.cc:172-178), at the NS state W_dev (its ghost slots are written).
 * opts (may be NULL: the reference's defaults, restart 30) are kept for the
 * solve that fcfd_newton_step runs.  Synchronises the device. */
int  fcfd_rans_init(fcfd_ctx* ctx, double* W_dev, double intensity, double visc_ratio, const fcfd_rans_opts* opts, void* stream);
/* k, omega of the interior cells from the host (N x 2, reference numbering),
 * then ghosts, eddy viscosity and gradients at W_dev (a restart's
 * read_restart, rans_read_restart.cc).  Needs fcfd_rans_init first (the inlet
 * values).  Synchronises the stream. */
int  fcfd_rans_set(fcfd_ctx* ctx, double* W_dev, const double* K_host, void* stream);
/* To the host in reference numbering (any may be NULL): K_host (N+B) x 2 =
 * k, omega; mut_host N+B; G_host N x 6 = grad k, grad omega; strain_host N+B
 * (the strain rate of the last terms evaluation); mutf_host F = mu_t.face.
 * Synchronises the device. */
int  fcfd_rans_get(fcfd_ctx* ctx, double* K_host, double* mut_host, double* G_host, double* strain_host, double* mutf_host);
/* RANS::terms + time_terms (rans_terms.cc:39-292, rans_time_terms.cc:25-86,
 * no pseudo time) at the NS state W_dev with the NS time step dt_dev ([ld]),
 * the NS face loop's mdot and weightL recomputed at W_dev: to the host in
 * reference numbering, rhs_host N x 2, diag_host N x 4 (row-major 2 x 2
 * diagonal blocks: faces, sources, rho V / dt), jf_host F x 6 (per face: its
 * rhs contributions to the parent, k and omega, then jacL_k, jacL_omega,
 * jacR_k, jacR_omega; the operator's row P, column N entry is jacR, row N,
 * column P is -jacL).  Any output may be NULL.  Recomputes ghosts, gradients
 * and limiters of W_dev; sets the strain rate.  Synchronises the stream. */
int  fcfd_rans_terms(fcfd_ctx* ctx, double* W_dev, const double* dt_dev, double* rhs_host, double* diag_host, double* jf_host,
                     void* stream);
/* RANS::solve (rans.cc:85-98) at the NS state W_dev with the NS time step
 * dt_dev: terms and time terms, GMRES with 2 x 2 block-Jacobi preconditioning
 * (PETSc FGMRES in the reference, rans_petsc_functions.cc:25-122),
 * update_variables with its limits (rans.cc:180-252), then ghosts, eddy
 * viscosity and gradients.  W_dev is not changed (its ghost slots are
 * written).  FCFD_EDIVERGED on a NaN/Inf update (k, omega unchanged; the
 * reference exits).  fcfd_newton_step runs the same solve between its
 * Krylov solve and its update, as NavierStokes::solve does (ns.cc:110). */
int  fcfd_rans_step(fcfd_ctx* ctx, double* W_dev, const double* dt_dev, const fcfd_rans_opts* opts, fcfd_rans_stats* stats,
                    void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FCFD_GPU_H */
