// Launchers of the adjoint (VJP) and surface-load kernels (fcfd_adjoint.hip).
#pragma once
#include "fcfd_device.h"

namespace fcfd {

// Connectivity the reverse sweep gathers over (built once per context).
struct AdjGrid {
    const int* hf_twin;   // nhf: half-face of the same face in the other cell, -1 for a boundary face
    const int* gh_hf;     // B: the parent's half-face of each ghost's face
    const int* tg_ptr;    // N+B+1: transposed gradient map, by stencil member j ...
    const int* tg_cell;   // ... the cells c whose gradient reads j, ascending
    const double *tgx, *tgy, *tgz;
    const int* tgm;       // per entry, the variables a BC holds fixed on that face (bit v); null without Green-Gauss cells
    // surface loads: per BC, the boundary faces in reference face order
    const int* lf_ptr;    // n_bc+1
    const int* lf_face;   // internal face ids
    int n_bc;
};

// Workspace of one reverse sweep.
struct AdjWork {
    double* Z;     // [32][F] adjoints of the face inputs (face states, centre states, face gradients)
    double* Wb1;   // [5][ld]  direct state adjoints (flux + own limiter), ghosts included
    double* Gb;    // [15][ld] gradient adjoints
    double* LH;    // [5][nhf] limiter adjoint towards the neighbour across each half-face
    double* Wb2;   // [5][ld]  after the transposed gradient map and limiter gathers
};

// Wbar = (dR/dW)^T Rbar + d/dW sum_{boundary faces f} fxw[bc(f)] . (D - C)_f, the exact
// transpose of launch_jvp(full = true).  Runs the primal ghosts, gradients and
// limiters first (writes W's ghost slots, G, PHI).  fxw: device [n_bc][5] or null.
void launch_vjp(const DevGrid& g, const AdjGrid& a, const Params& P, double* W, double* G, double* PHI, double* dref,
                const double* Rbar, const double* fxw, const AdjWork& w, double* Wbar, int* status, hipStream_t s);

// Surface loads (force, moment about mc) per BC into out[6 * n_bc] (device),
// each summed in reference face order over the BC's faces listed in
// lf_ptr/lf_face (empty for a BC not included in the loads); tmp is [6][nlf]
// scratch.  G/PHI must hold W's primal gradients and limiters.
void launch_loads(const DevGrid& g, const AdjGrid& a, const Params& P, const double* W, const double* G, const double* PHI,
                  double mcx, double mcy, double mcz, double* tmp, int nlf, double* out, int* status, hipStream_t s);
// the loads' tangent from the faces' d(D - C) (the same ordered sums)
void launch_loads_tangent(const DevGrid& g, const AdjGrid& a, const double* dFX, double mcx, double mcy, double mcz, double* tmp,
                          int nlf, double* out, hipStream_t s);
// out[0] = sum over boundary faces of sum_{v<5} w[5 bc + v] dFX[v][f]
// (part: 256 doubles)
void launch_bface_dot(const DevGrid& g, const double* w, const double* dFX, double* part, double* out, hipStream_t s);
}  // namespace fcfd
