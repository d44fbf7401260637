// Residual kernels for gfx950 (FP64).
//
// Three launches per residual evaluation, all deterministic and atomic-free:
//   1. ghost_kernel      one thread per boundary face: boundary ghost state
//   2. grad_lim_kernel   one thread per cell: gradient-map gather (sliced
//                        ELL, 64-wide slices matching the wavefront) fused
//                        with the limiter; also copies gradients to ghosts
//   3. flux_kernel       one thread per half-face (cell, face) pair: each
//                        block owns whole cells, computes the face fluxes of
//                        their faces, stages them in LDS and sums every cell's
//                        contributions in reference face order.  An interior
//                        face is evaluated by both of its cells with
//                        identical inputs and operation order, so both sides
//                        see the bitwise-same flux and R is conservative.
#include <hip/hip_runtime.h>

#include "fcfd_device.h"
#include "fcfd_face.h"

namespace fcfd {

// --------------------------------------------------------------------------
// 1. boundary ghosts: apply_bcs with the parent's cell-centre state as left and
//    right pre-set to a copy of left; ghost = right.{p,V,T}_center
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) ghost_kernel(DevGrid g, Params P, double* __restrict__ W, int* __restrict__ status) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.B) return;
    int f = g.gh_face[i];
    int c = g.fpar[f];
    FaceGeo F;
    load_geo(g, f, F);
    const size_t ld = g.ld;
    CState<double> L;
    L.p = L.p_center = W[c];
    L.V = L.V_center = mk<double>(W[ld + c], W[2 * ld + c], W[3 * ld + c]);
    L.T = L.T_center = W[4 * ld + c];
    L.rho = m_rho(P, L.p, L.T);
    finish_state(P, L, F);
    CState<double> R = L;
    int flag = apply_bcs(P, g.bc[F.bc], L, R, F);
    if (flag) atomicOr(status, 1);
    int gi = g.N + i;
    W[gi] = R.p_center;
    W[ld + gi] = R.V_center.x;
    W[2 * ld + gi] = R.V_center.y;
    W[3 * ld + gi] = R.V_center.z;
    W[4 * ld + gi] = R.T_center;
}

// --------------------------------------------------------------------------
// 2. gradients (calc_cell_grads, ns.cc:301-336 + variable.h:193-216) fused
//    with the limiter (ns_limiters.cc:26-248)
// --------------------------------------------------------------------------
template <int LIM>
__global__ void __launch_bounds__(256) grad_lim_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                       double* __restrict__ G, double* __restrict__ PHI,
                                                       const double* __restrict__ dref) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double gr[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) gr[k] = 0.0;
    {
        int s = c >> 6, lane = c & 63;
        int e0 = g.gslice[s], e1 = g.gslice[s + 1];
        for (int e = e0 + lane; e < e1; e += 64) {
            int j = g.gidx[e];
            double wx = g.gwx[e], wy = g.gwy[e], wz = g.gwz[e];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double q = W[v * ld + j];
                gr[3 * v] += wx * q;
                gr[3 * v + 1] += wy * q;
                gr[3 * v + 2] += wz * q;
            }
        }
    }
    const int h0 = g.hf_ptr[c], h1 = g.hf_ptr[c + 1];
#pragma unroll
    for (int k = 0; k < 15; ++k) G[k * ld + c] = gr[k];
    double q0[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) q0[v] = W[v * ld + c];
    double phi[5] = {1., 1., 1., 1., 1.};
    Vec<double> xc = centroid(g, c);
    if (LIM == FCFD_LIMITER_MINMOD) {
        const double small = 1.e-12;
        for (int h = h0; h < h1; ++h) {
            int fs = g.hf_face[h];
            int f = fs >= 0 ? fs : ~fs;
            int nb = fs >= 0 ? g.fnei[f] : g.fpar[f];
            Vec<double> c2f = vsub(mk<double>(g.fcx[f], g.fcy[f], g.fcz[f]), xc);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double dU = W[v * ld + nb] - q0[v];
                double dUg = gr[3 * v] * c2f.x + gr[3 * v + 1] * c2f.y + gr[3 * v + 2] * c2f.z;
                double mm = (dU * dUg < 0.) ? 0. : (dU > 0.) ? smin(dU, dUg) : smax(dU, dUg);
                double sU = (double)((mm > 0) - (mm < 0)), sG = (double)((dUg > 0) - (dUg < 0));
                phi[v] = smin(phi[v], (mm + sU * small) / (dUg + sG * small));
            }
        }
    } else if (LIM == FCFD_LIMITER_BJ || LIM == FCFD_LIMITER_VK) {
        double umax[5], umin[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) umax[v] = umin[v] = q0[v];
        for (int h = h0; h < h1; ++h) {
            int fs = g.hf_face[h];
            int f = fs >= 0 ? fs : ~fs;
            int nb = fs >= 0 ? g.fnei[f] : g.fpar[f];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double q = W[v * ld + nb];
                umax[v] = smax(umax[v], q);
                umin[v] = smin(umin[v], q);
            }
        }
        for (int h = h0; h < h1; ++h) {
            int fs = g.hf_face[h];
            int f = fs >= 0 ? fs : ~fs;
            Vec<double> c2f = vsub(mk<double>(g.fcx[f], g.fcy[f], g.fcz[f]), xc);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double dM = gr[3 * v] * c2f.x + gr[3 * v + 1] * c2f.y + gr[3 * v + 2] * c2f.z;
                double dP = q0[v];
                if (LIM == FCFD_LIMITER_BJ) {
                    if (dM > 0.) dP = umax[v] - dP;
                    else if (dM < 0.) dP = umin[v] - dP;
                    else { dP = 1.; dM = 1.; }
                    phi[v] = smin(phi[v], dP / dM);
                } else {
                    double eps = dref[v] / P.glscale;
                    if (dM > 0.) dP = umax[v] - dP;
                    else if (dM < 0.) dP = umin[v] - dP;
                    double dP2 = dP * dP, dM2 = dM * dM;
                    double e = 1.e-6 * P.limiter_threshold * eps;
                    double eps2 = e * e;
                    if (dM != 0.) phi[v] = smin(phi[v], ((dP2 + eps2) + 2. * dM * dP) / (dP2 + 2. * dM2 + dM * dP + eps2));
                }
            }
        }
    }
    if (LIM != FCFD_LIMITER_NONE) {
#pragma unroll
        for (int v = 0; v < 5; ++v) PHI[v * ld + c] = phi[v];
    }
    // boundary ghosts take their parent's gradient (ns.cc:320-333)
    for (int h = h0; h < h1; ++h) {
        int fs = g.hf_face[h];
        if (fs < 0) continue;
        int gi = g.fnei[fs];
        if (gi < g.N) continue;
#pragma unroll
        for (int k = 0; k < 15; ++k) G[k * ld + gi] = gr[k];
    }
}

// find_min_max (ns.cc:400-435) with the reference's index slips, for VK's
// reference range.  The reference loop sets qmax[3] = max(qmax[4], w_c) and
// then qmax[4] = max(qmax[4], p_c), with qmax[4] seeded by T of cell 0.  After
// the loop: slot 4 = max(T_0, max_c p_c); slot 3 = max(max(T_0, max_{c<last}
// p_c), w_last) where `last` is the last cell in reference order.  Min
// likewise.  `first`/`last` are the internal indices of reference cells 0 and
// N-1.  Two passes: fixed-size grid of partial min/max, then one block (max and
// min are exact, so the result is independent of the order anyway).
constexpr int kMMBlocks = 1024;
__global__ void __launch_bounds__(256) minmax_partial_kernel(int N, size_t ld, int first, int last, const double* __restrict__ W,
                                                             double* __restrict__ part) {
    __shared__ double sh[8][256];
    int t = threadIdx.x;
    // 0..2: max p,u,v   3: max p excluding `last`   4..6: min p,u,v   7: min p excluding `last`
    double r[8];
    r[0] = r[4] = W[first]; r[1] = r[5] = W[ld + first]; r[2] = r[6] = W[2 * ld + first];
    r[3] = r[7] = W[4 * ld + first];   // T_0 seeds the p-tracking slots
    for (int c = blockIdx.x * 256 + t; c < N; c += kMMBlocks * 256) {
        double p = W[c], u = W[ld + c], v = W[2 * ld + c];
        r[0] = smax(r[0], p); r[1] = smax(r[1], u); r[2] = smax(r[2], v);
        r[4] = smin(r[4], p); r[5] = smin(r[5], u); r[6] = smin(r[6], v);
        if (c != last) { r[3] = smax(r[3], p); r[7] = smin(r[7], p); }
    }
    for (int k = 0; k < 8; ++k) sh[k][t] = r[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 4; ++k) sh[k][t] = smax(sh[k][t], sh[k][t + s]);
            for (int k = 4; k < 8; ++k) sh[k][t] = smin(sh[k][t], sh[k][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) for (int k = 0; k < 8; ++k) part[k * kMMBlocks + blockIdx.x] = sh[k][0];
}

__global__ void __launch_bounds__(256) minmax_final_kernel(size_t ld, int first, int last, const double* __restrict__ W,
                                                           const double* __restrict__ part, double* __restrict__ dref) {
    __shared__ double sh[8][256];
    int t = threadIdx.x;
    double r[8];
    for (int k = 0; k < 8; ++k) r[k] = part[k * kMMBlocks + t];
    for (int i = t + 256; i < kMMBlocks; i += 256)
        for (int k = 0; k < 8; ++k) r[k] = k < 4 ? smax(r[k], part[k * kMMBlocks + i]) : smin(r[k], part[k * kMMBlocks + i]);
    for (int k = 0; k < 8; ++k) sh[k][t] = r[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 4; ++k) sh[k][t] = smax(sh[k][t], sh[k][t + s]);
            for (int k = 4; k < 8; ++k) sh[k][t] = smin(sh[k][t], sh[k][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        double T0 = W[4 * ld + first];
        double pl = W[last], wl = W[3 * ld + last];
        double q4max = smax(smax(sh[3][0], pl), T0), q4min = smin(smin(sh[7][0], pl), T0);
        double q3max = smax(sh[3][0], wl), q3min = smin(sh[7][0], wl);
        dref[0] = fabs(sh[0][0] - sh[4][0]);
        dref[1] = fabs(sh[1][0] - sh[5][0]);
        dref[2] = fabs(sh[2][0] - sh[6][0]);
        dref[3] = fabs(q3max - q3min);
        dref[4] = fabs(q4max - q4min);
    }
}

// --------------------------------------------------------------------------
// 3. face fluxes + per-cell ordered accumulation (ns_assemble_linear_system.cc:75-140)
// --------------------------------------------------------------------------
template <int FLUX, int VISC>
__global__ void __launch_bounds__(256) face_flux_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                        const double* __restrict__ G, const double* __restrict__ PHI,
                                                        int has_phi, double phi_const, double* __restrict__ FX,
                                                        int* __restrict__ status) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    FaceCtx<double> X;
    CellLoader ld_{W, G, PHI, has_phi != 0, phi_const, (size_t)g.ld};
    face_base<FLUX, VISC>(g, P, ld_, f, X, status);
    const size_t Fs = g.F;
#pragma unroll
    for (int v = 0; v < 5; ++v) FX[v * Fs + f] = X.D[v] - X.C[v];
}

// R_c = sum over the faces of c in reference face order of (D-C)+0 (parent)
// or -1*(D-C)+0 (neighbour) -- the VecSetValues(ADD) sequence of
// ns_assemble_linear_system.cc:131-140 starting from VecSet(rhs, 0)
__global__ void __launch_bounds__(256) cell_sum_kernel(DevGrid g, const double* __restrict__ FX, double* __restrict__ Rout) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld, Fs = g.F;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int h = g.hf_ptr[c]; h < g.hf_ptr[c + 1]; ++h) {
        int fs = g.hf_face[h];
        if (fs >= 0) {
#pragma unroll
            for (int v = 0; v < 5; ++v) s[v] += FX[v * Fs + fs] + 0.;
        } else {
            int f = ~fs;
#pragma unroll
            for (int v = 0; v < 5; ++v) s[v] += -1. * FX[v * Fs + f] + 0.;
        }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) Rout[v * ld + c] = s[v];
}

// --------------------------------------------------------------------------
// local time step (time_step.cc:193-220) for CFL_LOCAL; CFL_MAX uses a
// deterministic min-reduction
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) dt_local_kernel(DevGrid g, Params P, const double* __restrict__ W, double cfl,
                                                       double* __restrict__ dt) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double p = W[c], T = W[4 * ld + c];
    Vec<double> V = mk<double>(W[ld + c], W[2 * ld + c], W[3 * ld + c]);
    double a = m_a(P, p, T);
    dt[c] = cfl * g.lscale[c] / (vmag(V) + a);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <int FLUX, int VISC>
static void launch_flux_t(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI,
                          int has_phi, double phi_const, double* FX, int* status, hipStream_t s) {
    hipLaunchKernelGGL((face_flux_kernel<FLUX, VISC>), dim3((g.F + 255) / 256), dim3(256), 0, s, g, P, W, G, PHI, has_phi,
                       phi_const, FX, status);
}

template <int FLUX>
static void launch_flux_v(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI,
                          int has_phi, double phi_const, double* FX, int* status, hipStream_t s) {
    if (P.viscous) launch_flux_t<FLUX, 1>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s);
    else launch_flux_t<FLUX, 0>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s);
}

void launch_ghosts(const DevGrid& g, const Params& P, double* W, int* status, hipStream_t s) {
    if (g.B > 0) hipLaunchKernelGGL(ghost_kernel, dim3((g.B + 255) / 256), dim3(256), 0, s, g, P, W, status);
}

void launch_grad_lim(const DevGrid& g, const Params& P, const double* W, double* G, double* PHI, double* dref, hipStream_t s) {
    int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    dim3 grid((g.N + 255) / 256), blk(256);
    if (lim == FCFD_LIMITER_VK) {
        // eps^2 = (1e-6 K dref/L)^2 vanishes for K = 0 (the default threshold,
        // read_inputs.cc:129) whatever dref is, provided L > 0: skip the reduction
        if (!(P.limiter_threshold == 0.0 && P.glscale > 0.0)) {
            hipLaunchKernelGGL(minmax_partial_kernel, dim3(kMMBlocks), dim3(256), 0, s, g.N, (size_t)g.ld, g.ref_first, g.ref_last, W,
                               dref + 8);
            hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, s, (size_t)g.ld, g.ref_first, g.ref_last, W, dref + 8, dref);
        }
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_VK>), grid, blk, 0, s, g, P, W, G, PHI, dref);
    } else if (lim == FCFD_LIMITER_BJ) {
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_BJ>), grid, blk, 0, s, g, P, W, G, PHI, dref);
    } else if (lim == FCFD_LIMITER_MINMOD) {
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_MINMOD>), grid, blk, 0, s, g, P, W, G, PHI, dref);
    } else {
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_NONE>), grid, blk, 0, s, g, P, W, G, PHI, dref);
    }
}

void launch_vk_range(const DevGrid& g, const Params& P, const double* W, double* dref, hipStream_t s) {
    int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    if (lim != FCFD_LIMITER_VK || (P.limiter_threshold == 0.0 && P.glscale > 0.0)) return;
    hipLaunchKernelGGL(minmax_partial_kernel, dim3(kMMBlocks), dim3(256), 0, s, g.N, (size_t)g.ld, g.ref_first, g.ref_last, W, dref + 8);
    hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, s, (size_t)g.ld, g.ref_first, g.ref_last, W, dref + 8, dref);
}

void launch_flux(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, double* FX,
                 double* R, int* status, hipStream_t s) {
    // limiter buffer is used only when a limiter runs; otherwise phi is the
    // constant initial value (ns.cc:282-287): 1 for second order, 0 for first
    int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    int has_phi = lim != FCFD_LIMITER_NONE;
    double phi_const = (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0;
    switch (P.flux) {
    case FCFD_FLUX_ROE: launch_flux_v<FCFD_FLUX_ROE>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s); break;
    case FCFD_FLUX_VAN_LEER: launch_flux_v<FCFD_FLUX_VAN_LEER>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s); break;
    case FCFD_FLUX_SD_SLAU: launch_flux_v<FCFD_FLUX_SD_SLAU>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s); break;
    case FCFD_FLUX_SW: launch_flux_v<FCFD_FLUX_SW>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s); break;
    default: launch_flux_v<FCFD_FLUX_AUSM_PLUS_UP>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s); break;
    }
}

void launch_cell_sum(const DevGrid& g, const double* FX, double* R, hipStream_t s) {
    hipLaunchKernelGGL(cell_sum_kernel, dim3((g.N + 255) / 256), dim3(256), 0, s, g, FX, R);
}

void launch_dt_local(const DevGrid& g, const Params& P, const double* W, double cfl, double* dt, hipStream_t s) {
    hipLaunchKernelGGL(dt_local_kernel, dim3((g.N + 255) / 256), dim3(256), 0, s, g, P, W, cfl, dt);
}

}  // namespace fcfd
