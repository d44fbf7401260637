// Launchers of the tangent (JVP) kernels (fcfd_tangent.hip).
#pragma once
#include "fcfd_device.h"

namespace fcfd {
// dR = (dR/dW) dW.  full: differentiate ghosts, gradients and limiters too
// (writes the ghost slots of W and dW, G/dG and PHI/dPHI); otherwise the
// frozen-gradient linearisation.  dFX is [5][F] scratch.
void launch_jvp(const DevGrid& g, const Params& P, double* W, double* dW, double* G, double* dG, double* PHI, double* dPHI,
                double* dref, double* dFX, double* dR, bool full, int* status, hipStream_t s);
// y = (P V / dt) x - dR
void launch_newton_op(const DevGrid& g, const Params& P, const double* W, const double* dt, const double* x, const double* dR,
                      double* y, hipStream_t s);
// y = (P V / dt)^T x - Jt   (Jt = (dR/dW)^T x)
void launch_newton_op_t(const DevGrid& g, const Params& P, const double* W, const double* dt, const double* x, const double* Jt,
                        double* y, hipStream_t s);
}
