// Segregated two-equation RANS kernels (fcfd_rans.hip).
#pragma once
#include "fcfd_device.h"

namespace fcfd {

// the model's per-cell fields (device)
struct KwDev {
    double* K;        // [2][ld] k, omega (interior cells and ghosts)
    double* mut;      // [ld] eddy viscosity (interior cells and ghosts)
    double* GK;       // [6][ld] grad k, grad omega (interior cells)
    double* strain;   // [ld] sqrt(2 S_ij S_ij) of the last terms evaluation
};
// one terms evaluation's outputs (device)
struct KwTerms {
    double* FV;         // [6][F] per-face parent contributions
    double *JL, *JR;    // 2 x 2 off-diagonal blocks per face (row pairs, the NS operator's convention)
    double* rhs;        // [2][ld]
    double* D;          // 2 x 2 diagonal blocks (row pairs, [2][ld] double2)
    double* lam_last;   // 1: the face loop's last laminar viscosity
};

// apply_initial_conditions over the whole domain (interior k, omega; gradients 0; strain 0)
void kw_launch_init(const DevGrid& g, const Params& P, const double* W, double intensity, double visc_ratio, const KwDev& k,
                    hipStream_t s);
// set_bcs: out[3 b .. 3 b + 2] = k, omega, mu_t of each inlet BC from the NS
// face values of its first face f0[b] (internal face id, -1 for none)
void kw_launch_inlets(const DevGrid& g, const Params& P, int n_bc, const double* W, const int* f0, double* out, hipStream_t s);
// ghosts, eddy viscosity, gradients (after init, set or update)
void kw_launch_refresh(const DevGrid& g, const Params& P, const double* W, const KwDev& k, hipStream_t s);
// RANS::terms + time_terms at the NS state W (its ghosts, gradients G and
// limiters PHI current) with the NS time step dt; last_face = internal id of
// the reference's last face
void kw_launch_terms(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, const double* dt,
                     const KwDev& k, const KwTerms& t, int last_face, int* status, hipStream_t s);
// update_variables: out_host = sum dk^2/dt^2, sum domega^2/dt^2 (limited
// updates), NaN/Inf count of dU; k, omega += the limited update only when
// the count is 0.  part: 3 * 256 doubles, out_dev: 3.
hipError_t kw_update(const DevGrid& g, const Params& P, const KwDev& k, const double* dU, const double* dt, double* part,
                     double* out_dev, double out_host[3], hipStream_t s);

}  // namespace fcfd
