/*
 * fcfd_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64, no FMA contraction) of Free CFD's
 * residual/Jacobian path, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the parity checker.  Nothing in the
 * product (freecfd_amd/, include/) links or calls it.
 */
#ifndef FCFD_ORACLE_H
#define FCFD_ORACLE_H
#include "../include/fcfd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* set-up: Grid::setup (create_nodes_cells, create_faces, areas_volumes,
 * create_boundary_ghosts), set_bcs, set_lengthScales,
 * face_interpolation_weights, gradient_maps.  Returns NULL on error; the
 * message is then in orc_last_error(NULL). */
orc_ctx* orc_create(const fcfd_mesh* mesh, const fcfd_options* opt);
void orc_destroy(orc_ctx* o);
const char* orc_last_error(const orc_ctx* o);
int orc_info(const orc_ctx* o, fcfd_info* info);
int orc_metrics(const orc_ctx* o, fcfd_metrics_out* m);

/* state W: N x nv (p,u,v,w,T[,nu~]; nv = 6 with the SA model), reference
 * numbering; every per-cell array below uses the same nv */
int orc_set_state(orc_ctx* o, const double* W);
int orc_set_update(orc_ctx* o, const double* dW);   /* previous Newton update (update[i].cell) */
int orc_set_full_state(orc_ctx* o, const double* W);  /* (N+B) x 5 incl. ghosts */
int orc_flux_probe(orc_ctx* o, int flux_type, const double* WL, const double* WR, const double* n, const double* t1,
                   double area, double* out);
int orc_update_boundaries(orc_ctx* o);
int orc_calc_cell_grads(orc_ctx* o);
int orc_calc_limiter(orc_ctx* o);
/* face loop; with_jacobian adds the FD Jacobian (get_jacobians) */
int orc_assemble(orc_ctx* o, int with_jacobian);
/* R, sum|contrib| : N x 5 */
int orc_get_residual(const orc_ctx* o, double* R, double* Rabs);
/* full W incl. ghosts (N+B) x 5, grads (N+B) x 15, limiter N x 5 */
int orc_get_fields(const orc_ctx* o, double* Wfull, double* G, double* phi);
/* FD Jacobian: per face JL,JR [F][5 var][5 eq]; diag blocks [N][5 row][5 col]
 * (face contributions only, before time terms) */
int orc_get_jacobian(const orc_ctx* o, double* JL, double* JR, double* Dg);
/* loads per BC: force xyz, moment xyz */
int orc_get_loads(const orc_ctx* o, double* loads);
/* local time step, CFL_LOCAL or CFL_MAX (time_step.cc:143-224) */
int orc_time_step(orc_ctx* o, int dt_type, double cfl, double* dt, double* max_cfl);
/* time_terms P*V/dt blocks [N][5][5] (ns_time_terms.cc:26-56,147-163) */
int orc_time_terms(orc_ctx* o, const double* dt, double* blocks);
/* per-face convective C and diffusive D fluxes of the last orc_assemble, F x nv each */
int orc_get_face_flux(const orc_ctx* o, double* C, double* D);
/* one implicit step at CFL_LOCAL on the host cores (FD Jacobian, time terms,
 * block-Jacobi GMRES(restart), W += dW); stats[6] = ||R||_2, Krylov
 * iterations, ms of residual+Jacobian, time terms+inverses, Krylov, update */
int orc_newton_step(orc_ctx* o, double cfl, int restart, int max_iters, double rtol, double* stats);
int orc_num_threads(void);
/* 1: wall distances by the reference's brute-force scan in later orc_create
 * calls; 0 (default): the exact bucketed search, bitwise the same */
int orc_set_wall_brute_force(int on);
/* one pseudo step's system after orc_assemble: blocks = P V/dt + P_ps V/dtau
 * (P_ps = cons2prim, or WS95 with ws95), rhs = R - (P V/dt)(W - Wn) for
 * ps_step > 1 (ns_time_terms.cc:58-186); dt, dtau N; Wn, rhs N x nv */
int orc_pseudo_system(orc_ctx* o, const double* dt, const double* dtau, int ws95, int ps_step, const double* Wn, double* blocks,
                      double* rhs);
/* mdot.face, qdot.bc, tau.bc of the last orc_assemble (F, F, 3F; reference face order) */
int orc_get_surface(const orc_ctx* o, double* mdot, double* qdot, double* tau);
/* apply_bcs on BC `bc`: left face state WL and centre WLc, ghost centre WRc
 * (p,u,v,w,T each), face normal n and tangent t1; out (12) = right state
 * p,u,v,w,T,rho, the a apply_bcs set, centre p,u,v,w,T.  FCFD_EBC where the
 * reference exits. */
int orc_bc_probe(orc_ctx* o, int bc, const double* WL, const double* WLc, const double* WRc, const double* n,
                 const double* t1, double* out);
/* gelimd (utilities.cc:71-101) on an n x n row-major system, n <= 4 */
int orc_gelimd_probe(int n, const double* a, const double* b, double* x);
/* Interpolate::calculate_weights(is_internal = true) on one stencil (fresh
 * object): kept ids and weights, *kind 4/3/2/1 = tetra/tri/line/point; returns
 * the number of weights */
int orc_wtli_probe(int dim, int maxs, double tol, const double* point, int n, const int* ids, const double* pts,
                   int* out_ids, double* out_w, int* kind);
/* AUSM+-up split functions: out = M4+(M), M4-(M), P5+(M,alpha), P5-(M,alpha) */
int orc_ausm_probe(double M, double alpha, double* out);
/* SA model pieces at one point: out = source/volume, P, D, mu_t, fn */
int orc_sa_probe(double rho, double mu, double nut, const double* gvel, const double* gn, double d, double* out);

/* segregated two-equation RANS (src/rans/), for opt.turbulence = FCFD_TURB_KW.
 * The NS state, its ghosts and gradients must be current (orc_set_state,
 * orc_update_boundaries, orc_calc_cell_grads) before init/set; orc_assemble
 * at the same state before orc_kw_terms (mdot and weightL of its faces). */
int orc_kw_init(orc_ctx* o, double intensity, double visc_ratio);
int orc_kw_set(orc_ctx* o, const double* K);
int orc_kw_bc_values(const orc_ctx* o, int b, double* out);
int orc_kw_terms(orc_ctx* o, const double* dt, double* rhs, double* diag, double* jf);
int orc_kw_update(orc_ctx* o, const double* dU, const double* dt, double* sums);
int orc_kw_get(const orc_ctx* o, double* K, double* mut, double* G, double* strain, double* mutf);
int orc_get_weight_left(const orc_ctx* o, double* wl);
int orc_kw_source_probe(int model, double rho, double k, double omega, double mu_t, double lam_visc, const double* gvel,
                        const double* gk, const double* go, double d, double vol, double* out);

#ifdef __cplusplus
}
#endif
#endif
