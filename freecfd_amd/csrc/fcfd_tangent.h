// Launchers of the tangent (JVP) kernels (fcfd_tangent.hip).
#pragma once
#include "fcfd_device.h"

namespace fcfd {
// The limiters' linearisation at a fixed state, for the FP32 operator of a
// Newton step (W fixed over the Krylov solve): per cell c and variable v,
// phi = min(1, min over c's faces of T_f) has one active term (or none), and
//   dphi_v = a (dq_nb - dq_c) + gv . dgrad_v
// with nb the neighbour whose state enters it (-1: none).  nb, a: [nv][ld];
// gv: [3 nv][ld].  Formed once per step (launch_lim_lin) by the operations
// and branch choices of the limiter itself, in FP64, stored in FP32.
struct LimLin {
    int* nb;
    float* a;
    float* gv;
};
void launch_lim_lin(const DevGrid& g, const Params& P, const double* W, const double* G, const double* dref, const LimLin& L,
                    hipStream_t s);
// dR = (dR/dW) dW.  full: differentiate ghosts, gradients and limiters too
// (writes the ghost slots of W and dW, G/dG and PHI/dPHI); otherwise the
// frozen-gradient linearisation.  dFX is [5][F] scratch.  primal_known: the
// ghost slots of W, G, PHI (and dref) already hold the residual pass's values
// at this W, so the primal gradients and limiters are read, not recomputed
// (the result is the same bit for bit; the Newton operator's case, where W
// is fixed over the Krylov solve).  dsrc (N entries, SA model, full): the
// source term's tangent is formed in the gradient kernel, kept there and
// added by the ordered cell sum (the same result as the separate source
// kernel used when dsrc is null).
void launch_jvp(const DevGrid& g, const Params& P, double* W, double* dW, double* G, double* dG, double* PHI, double* dPHI,
                double* dref, double* dFX, double* dR, bool full, int* status, hipStream_t s, bool primal_known = false,
                double* dsrc = nullptr);
// the same with an FP32 derivative lane: FP64 values (the residual's), FP32
// tangents dW, dG, dPHI, dFX, dR (not on meshes with Green-Gauss cells)
// lin (full, limited second order): the limiters' linearisation at W
// replaces the dual-number limiter in the gradient kernel; with primal_known
// it must be that of the step's W, otherwise it is recomputed here.
void launch_jvp(const DevGrid& g, const Params& P, double* W, float* dW, double* G, float* dG, double* PHI, float* dPHI,
                double* dref, float* dFX, float* dR, bool full, int* status, hipStream_t s, bool primal_known = false,
                float* dsrc = nullptr, const LimLin* lin = nullptr);
// y = (P V / dt) x - dR
void launch_newton_op(const DevGrid& g, const Params& P, const double* W, const double* dt, const double* x, const double* dR,
                      double* y, hipStream_t s);
void launch_newton_op(const DevGrid& g, const Params& P, const double* W, const double* dt, const float* x, const float* dR,
                      float* y, hipStream_t s);
// dual time stepping: drdp[c] = drho/dp of the pseudo block (cons2prim, or
// WS95 with ws95; nbp/nbi the interior cells' neighbour cells, read with
// ws95 only); with corr also R -= (P V/dt)(W - Wn)
void launch_pseudo_rhs(const DevGrid& g, const Params& P, const double* W, const double* Wn, const double* dt, const int* nbp,
                       const int* nbi, int ws95, int corr, double* R, double* drdp, hipStream_t s);
// D += P_ps V/dtau (diagonal blocks in row pairs)
void launch_pseudo_diag(const DevGrid& g, const Params& P, const double* W, const double* dtau, const double* drdp, double* D,
                        hipStream_t s);
// y += P_ps V/dtau x
void launch_pseudo_op(const DevGrid& g, const Params& P, const double* W, const double* dtau, const double* drdp, const double* x,
                      double* y, hipStream_t s);
void launch_pseudo_op(const DevGrid& g, const Params& P, const double* W, const double* dtau, const double* drdp, const float* x,
                      float* y, hipStream_t s);
// y = (P V / dt)^T x - Jt   (Jt = (dR/dW)^T x)
void launch_newton_op_t(const DevGrid& g, const Params& P, const double* W, const double* dt, const double* x, const double* Jt,
                        double* y, hipStream_t s);
// the parameter tangent: dR (and the faces' d(D - C) in dFX) at fixed W in
// the direction P.seed of one BC's values; dW [nv][ld] must be 0 on interior
// cells (its ghost rows, G, dG, PHI, dPHI are written)
void launch_param_tangent(const DevGrid& g, const Params& P, double* W, double* dW, double* G, double* dG, double* PHI, double* dPHI,
                          double* dref, double* dFX, double* dR, int* status, hipStream_t s);
}
