// Limiter of one cell (ns_limiters.cc:45-248) on a generic scalar type, for
// the tangent and adjoint kernels.  The FP64 residual keeps its own fused
// copy in grad_lim_kernel; this one performs the same operations in the same
// order, so on dual numbers the value part equals the residual's limiter and
// the derivative part is the exact derivative of the active branch.
//
// `src.nb(h, nb, v)` returns variable v of the neighbour `nb` across
// half-face h (the only place neighbour states enter); q0 and gr are the
// cell's own state and gradient; dref is the VK reference range (held
// constant).  Loop-based (the neighbour loads of one half-face at a time):
// on dual numbers the round-based load schedule of the FP64 kernel needs
// more registers than it saves latency.
#pragma once
#include "fcfd_face.h"

namespace fcfd {

template <int LIM, int NV, typename S, typename Src>
__device__ __forceinline__ void limiter_phi(const DevGrid& g, const Params& P, const Src& src, int c, const S q0[NV],
                                            const S gr[3 * NV], const double* __restrict__ dref, S phi[NV]) {
#pragma unroll
    for (int v = 0; v < NV; ++v) phi[v] = S(1.);
    const int h0 = g.hf_ptr[c], h1 = g.hf_ptr[c + 1];
    Vec<double> xc = centroid(g, c);
    if (LIM == FCFD_LIMITER_MINMOD) {
        const double small = 1.e-12;
        for (int h = h0; h < h1; ++h) {
            int fs = g.hf_face[h];
            int f = fs >= 0 ? fs : ~fs;
            int nb = g.hf_nb[h];
            Vec<double> c2f = vsub(face_centroid(g, f), xc);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                S dU = src.nb(h, nb, v) - q0[v];
                S dUg = gr[3 * v] * c2f.x + gr[3 * v + 1] * c2f.y + gr[3 * v + 2] * c2f.z;
                S mm = (val(dU) * val(dUg) < 0.) ? S(0.) : (val(dU) > 0.) ? smin(dU, dUg) : smax(dU, dUg);
                double sU = (double)((val(mm) > 0) - (val(mm) < 0)), sG = (double)((val(dUg) > 0) - (val(dUg) < 0));
                phi[v] = smin(phi[v], (mm + sU * small) / (dUg + sG * small));
            }
        }
    } else if (LIM == FCFD_LIMITER_BJ || LIM == FCFD_LIMITER_VK) {
        S umax[NV], umin[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) umax[v] = umin[v] = q0[v];
        for (int h = h0; h < h1; ++h) {
            int nb = g.hf_nb[h];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                S q = src.nb(h, nb, v);
                umax[v] = smax(umax[v], q);
                umin[v] = smin(umin[v], q);
            }
        }
        for (int h = h0; h < h1; ++h) {
            int fs = g.hf_face[h];
            int f = fs >= 0 ? fs : ~fs;
            Vec<double> c2f = vsub(face_centroid(g, f), xc);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                S dM = gr[3 * v] * c2f.x + gr[3 * v + 1] * c2f.y + gr[3 * v + 2] * c2f.z;
                S dP = q0[v];
                if (LIM == FCFD_LIMITER_BJ) {
                    if (val(dM) > 0.) dP = umax[v] - dP;
                    else if (val(dM) < 0.) dP = umin[v] - dP;
                    else { dP = S(1.); dM = S(1.); }
                    phi[v] = smin(phi[v], dP / dM);
                } else {
                    double eps = dref[v] / P.glscale;
                    if (val(dM) > 0.) dP = umax[v] - dP;
                    else if (val(dM) < 0.) dP = umin[v] - dP;
                    S dP2 = dP * dP, dM2 = dM * dM;
                    double e = 1.e-6 * P.limiter_threshold * eps;
                    double eps2 = e * e;
                    if (val(dM) != 0.) phi[v] = smin(phi[v], ((dP2 + eps2) + 2. * dM * dP) / (dP2 + 2. * dM2 + dM * dP + eps2));
                }
            }
        }
    }
}

}  // namespace fcfd
