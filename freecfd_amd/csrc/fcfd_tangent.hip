// Tangent (JVP) of the residual: dR = (dR/dW) dW, by forward-mode dual
// numbers through the same code as the FP64 residual.
//
//   FULL   ghosts (update_boundaries) -> gradients -> limiters -> face
//          fluxes (-> SA source) are all differentiated: the exact
//          Jacobian-vector product of the second-order residual.
//   FROZEN only the cell-centre states of the two cells of each face (and of
//          the cell itself in the SA source) carry a tangent; gradients,
//          limiters and ghost centres are frozen.  This is the linearisation
//          the reference FD Jacobian approximates (get_jacobians perturbs the
//          reconstructed face states only, ns_assemble_linear_system.cc:197-267),
//          i.e. dR = sum_f JL dW_P + JR dW_N with exact derivatives instead of
//          one-sided differences.
// The VK global reference range (find_min_max) is held constant; with the
// default threshold K = 0 it does not enter the limiter at all.  Every kernel
// is templated on NV (5, or 6 with the SA model).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fcfd_face.h"
#include "fcfd_limiter.h"
#include "fcfd_tangent.h"

namespace fcfd {

// neighbour states with their tangents, for limiter_phi
template <typename TD>
struct TangentNb {
    const double* __restrict__ W;
    const TD* __restrict__ dW;
    size_t ld;
    __device__ __forceinline__ DualT<TD> nb(int, int j, int v) const { return DualT<TD>(W[v * ld + j], dW[v * ld + j]); }
};

// PSEED: the values of BC P.seed.index carry the parameter tangent's direction
template <int NV, typename TD, bool PSEED = false>
__global__ void __launch_bounds__(256) ghost_t_kernel(DevGrid g, Params P, double* __restrict__ W, TD* __restrict__ dW,
                                                      int* __restrict__ status) {
    using Dual = DualT<TD>;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.B) return;
    int f = g.gh_face[i];
    int c = g.fpar[f];
    FaceGeo F;
    load_geo(g, f, F);
    const size_t ld = g.ld;
    CState<Dual> L;
    L.p = L.p_center = Dual(W[c], dW[c]);
    L.V = L.V_center = mk<Dual>(Dual(W[ld + c], dW[ld + c]), Dual(W[2 * ld + c], dW[2 * ld + c]), Dual(W[3 * ld + c], dW[3 * ld + c]));
    L.T = L.T_center = Dual(W[4 * ld + c], dW[4 * ld + c]);
    if constexpr (NV == 6) L.nut = L.nut_center = Dual(W[5 * ld + c], dW[5 * ld + c]);
    L.rho = m_rho(P, L.p, L.T);
    finish_state(P, L, F);
    CState<Dual> R = L;
    int flag;
    if constexpr (PSEED) {
        if (F.bc == P.seed.index) flag = apply_bcs(P, bc_val<Dual>(g.bc[F.bc], P.seed), L, R, F);
        else flag = apply_bcs(P, g.bc[F.bc], L, R, F);
    } else {
        flag = apply_bcs(P, g.bc[F.bc], L, R, F);
    }
    if (flag) atomicOr(status, 1);
    int gi = g.N + i;
    W[gi] = R.p_center.v;          dW[gi] = R.p_center.d;
    W[ld + gi] = R.V_center.x.v;   dW[ld + gi] = R.V_center.x.d;
    W[2 * ld + gi] = R.V_center.y.v; dW[2 * ld + gi] = R.V_center.y.d;
    W[3 * ld + gi] = R.V_center.z.v; dW[3 * ld + gi] = R.V_center.z.d;
    W[4 * ld + gi] = R.T_center.v; dW[4 * ld + gi] = R.T_center.d;
    if constexpr (NV == 6) { W[5 * ld + gi] = R.nut_center.v; dW[5 * ld + gi] = R.nut_center.d; }
}

// The limiters' linearisation at W (LimLin): the FP64 limiter of
// limiter_phi (the same operations, neighbour order and tie rules: smin keeps
// the earlier value, smax likewise), tracking which face's term is active
// and, for it, the partial derivatives with respect to its two arguments:
//   BJ  T = dP / dM,                       dP = u_max|min - q_c, dM = grad.c2f
//   VK  T = (dP^2 + e^2 + 2 dM dP) / (dP^2 + 2 dM^2 + dM dP + e^2)
//   minmod T = (mm + s_mm small) / (dM + s_dM small), mm = 0, dU or dM
// dP's neighbour is the cell that set u_max (u_min), -1 when q_c itself did.
template <int LIM, int NV>
__global__ void __launch_bounds__(256) lim_lin_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ G,
                                                      const double* __restrict__ dref, LimLin L) {
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double q0[NV], gr[3 * NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) q0[v] = W[v * ld + c];
    load_pairs<3 * NV>(G, ld, c, gr);
    double phi[NV], a[NV], gx[NV], gy[NV], gz[NV];
    int nbA[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { phi[v] = 1.; a[v] = 0.; gx[v] = gy[v] = gz[v] = 0.; nbA[v] = -1; }
    const int h0 = g.hf_ptr[c], h1 = g.hf_ptr[c + 1];
    const Vec<double> xc = centroid(g, c);
    if (LIM == FCFD_LIMITER_MINMOD) {
        const double small = 1.e-12;
        for (int h = h0; h < h1; ++h) {
            const int fs = g.hf_face[h], f = fs >= 0 ? fs : ~fs, nb = g.hf_nb[h];
            const Vec<double> c2f = vsub(face_centroid(g, f), xc);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double dU = W[v * ld + nb] - q0[v];
                const double dUg = gr[3 * v] * c2f.x + gr[3 * v + 1] * c2f.y + gr[3 * v + 2] * c2f.z;
                int pick;   // 0: mm = 0, 1: mm = dU, 2: mm = dUg
                double mm;
                if (dU * dUg < 0.) { pick = 0; mm = 0.; }
                else if (dU > 0.) { pick = (dUg < dU) ? 2 : 1; mm = pick == 2 ? dUg : dU; }   // smin(dU, dUg)
                else { pick = (dU < dUg) ? 2 : 1; mm = pick == 2 ? dUg : dU; }                 // smax(dU, dUg)
                const double sU = (double)((mm > 0) - (mm < 0)), sG = (double)((dUg > 0) - (dUg < 0));
                const double den = dUg + sG * small;
                const double T = (mm + sU * small) / den;
                if (T < phi[v]) {
                    phi[v] = T;
                    double cm;
                    if (pick == 1) { nbA[v] = nb; a[v] = 1. / den; cm = -T / den; }
                    else { nbA[v] = -1; a[v] = 0.; cm = pick == 2 ? (1. - T) / den : -T / den; }
                    gx[v] = cm * c2f.x; gy[v] = cm * c2f.y; gz[v] = cm * c2f.z;
                }
            }
        }
    } else if (LIM == FCFD_LIMITER_BJ || LIM == FCFD_LIMITER_VK) {
        double umax[NV], umin[NV];
        int imax[NV], imin[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { umax[v] = umin[v] = q0[v]; imax[v] = imin[v] = -1; }
        for (int h = h0; h < h1; ++h) {
            const int nb = g.hf_nb[h];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double q = W[v * ld + nb];
                if (umax[v] < q) { umax[v] = q; imax[v] = nb; }
                if (q < umin[v]) { umin[v] = q; imin[v] = nb; }
            }
        }
        for (int h = h0; h < h1; ++h) {
            const int fs = g.hf_face[h], f = fs >= 0 ? fs : ~fs;
            const Vec<double> c2f = vsub(face_centroid(g, f), xc);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double dM = gr[3 * v] * c2f.x + gr[3 * v + 1] * c2f.y + gr[3 * v + 2] * c2f.z;
                double dP = q0[v];
                int nbc = -1;
                if (LIM == FCFD_LIMITER_BJ) {
                    bool flat = false;
                    if (dM > 0.) { dP = umax[v] - dP; nbc = imax[v]; }
                    else if (dM < 0.) { dP = umin[v] - dP; nbc = imin[v]; }
                    else { dP = 1.; dM = 1.; flat = true; }
                    const double T = dP / dM;
                    if (T < phi[v]) {
                        phi[v] = T;
                        const double cm = flat ? 0. : -T / dM;
                        nbA[v] = flat ? -1 : nbc;
                        a[v] = (flat || nbc < 0) ? 0. : 1. / dM;
                        gx[v] = cm * c2f.x; gy[v] = cm * c2f.y; gz[v] = cm * c2f.z;
                    }
                } else {
                    const double eps = dref[v] / P.glscale;
                    if (dM > 0.) { dP = umax[v] - dP; nbc = imax[v]; }
                    else if (dM < 0.) { dP = umin[v] - dP; nbc = imin[v]; }
                    const double dP2 = dP * dP, dM2 = dM * dM;
                    const double e = 1.e-6 * P.limiter_threshold * eps;
                    const double eps2 = e * e;
                    if (dM != 0.) {
                        const double D = dP2 + 2. * dM2 + dM * dP + eps2;
                        const double T = ((dP2 + eps2) + 2. * dM * dP) / D;
                        if (T < phi[v]) {
                            phi[v] = T;
                            const double cp = ((2. * dP + 2. * dM) - T * (2. * dP + dM)) / D;
                            const double cm = (2. * dP - T * (4. * dM + dP)) / D;
                            nbA[v] = nbc;
                            a[v] = nbc >= 0 ? cp : 0.;
                            gx[v] = cm * c2f.x; gy[v] = cm * c2f.y; gz[v] = cm * c2f.z;
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        L.nb[v * ld + c] = nbA[v];
        L.a[v * ld + c] = (float)a[v];
        L.gv[(3 * v) * ld + c] = (float)gx[v];
        L.gv[(3 * v + 1) * ld + c] = (float)gy[v];
        L.gv[(3 * v + 2) * ld + c] = (float)gz[v];
    }
}

void launch_lim_lin(const DevGrid& g, const Params& P, const double* W, const double* G, const double* dref, const LimLin& L,
                    hipStream_t s) {
    const int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    const dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
#define FCFD_LIM_LIN(LM)                                                                                                     \
    do {                                                                                                                     \
        if (g.nv == 6) hipLaunchKernelGGL((lim_lin_kernel<LM, 6>), grid, blk, 0, s, g, P, W, G, dref, L);                    \
        else hipLaunchKernelGGL((lim_lin_kernel<LM, 5>), grid, blk, 0, s, g, P, W, G, dref, L);                              \
    } while (0)
    if (lim == FCFD_LIMITER_VK) FCFD_LIM_LIN(FCFD_LIMITER_VK);
    else if (lim == FCFD_LIMITER_BJ) FCFD_LIM_LIN(FCFD_LIMITER_BJ);
    else if (lim == FCFD_LIMITER_MINMOD) FCFD_LIM_LIN(FCFD_LIMITER_MINMOD);
#undef FCFD_LIM_LIN
}

// KNOWN: G and PHI already hold the primal gradients and limiters of W (the
// residual pass at the same state wrote them, with the same arithmetic in
// the same order), so the primal gradient is loaded, not re-gathered, and
// neither G nor PHI is written.  The derivative lane is the same either way.
// LIN (with KNOWN, FP32 lanes): the limiters' derivative from their
// linearisation at W (LimLin) instead of the limiter on dual numbers.
template <int LIM, int NV, typename TD, bool KNOWN, bool LIN = false>
__global__ void __launch_bounds__(256) grad_lim_t_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                         const TD* __restrict__ dW, double* __restrict__ G,
                                                         TD* __restrict__ dG, double* __restrict__ PHI,
                                                         TD* __restrict__ dPHI, const double* __restrict__ dref,
                                                         TD* __restrict__ dSRC, LimLin lin) {
    using Dual = DualT<TD>;
    int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double gv[3 * NV];
    TD gd[3 * NV];
#pragma unroll
    for (int k = 0; k < 3 * NV; ++k) { gv[k] = 0.0; gd[k] = TD(0); }
    {   // map entries in chunks: their weights in one round, their states in the next
        constexpr int kChunk = 3;
        const int s = c >> 6, lane = c & 63;
        const int e0 = g.gslice[s], K = (g.gslice[s + 1] - e0) >> 6;
        for (int k0 = 0; k0 < K; k0 += kChunk) {
            double2 a[kChunk], b[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int e = e0 + 64 * min(k0 + u, K - 1) + lane;
                a[u] = g.gwa[e];
                b[u] = g.gwb[e];
            }
            Dual q[kChunk][NV];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int j = (int)__double_as_longlong(b[u].y);
#pragma unroll
                for (int v = 0; v < NV; ++v) q[u][v] = Dual(KNOWN ? 0.0 : W[v * ld + j], dW[v * ld + j]);
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                if (k0 + u < K) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        // the dual sums gr += w q: value and derivative lanes
                        // apart (the value lane is dropped when KNOWN)
                        const Dual x = a[u].x * q[u][v], y = a[u].y * q[u][v], z = b[u].x * q[u][v];
                        if constexpr (!KNOWN) {
                            gv[3 * v] += x.v;
                            gv[3 * v + 1] += y.v;
                            gv[3 * v + 2] += z.v;
                        }
                        gd[3 * v] += x.d;
                        gd[3 * v + 1] += y.d;
                        gd[3 * v + 2] += z.d;
                    }
                }
            }
        }
    }
    const int h0 = g.hf_ptr[c], nh = g.hf_ptr[c + 1] - h0;
    if (KNOWN || (g.ggc && g.ggc[c])) load_pairs<3 * NV>(G, ld, c, gv);
    if (g.ggc && g.ggc[c]) load_pairs<3 * NV>(dG, ld, c, gd);   // a Green-Gauss cell: launch_green_gauss's tangent
    Dual gr[3 * NV];
#pragma unroll
    for (int k = 0; k < 3 * NV; ++k) gr[k] = Dual(gv[k], gd[k]);
    if constexpr (!KNOWN) store_pairs<3 * NV>(G, ld, c, gv);
    store_pairs<3 * NV>(dG, ld, c, gd);
    if constexpr (NV == 6) {
        // the tangent of V_c S_c (SA source), summed into the nu~ row by
        // cell_sum after the faces: sa_source_t_kernel's arithmetic on the
        // same duals
        if (dSRC) {
            const Dual p(W[c], dW[c]), T(W[4 * ld + c], dW[4 * ld + c]), nut(W[5 * ld + c], dW[5 * ld + c]);
            auto gvec = [&](int v) {
                Vec<Dual> r;
                r.x = gr[3 * v];
                r.y = gr[3 * v + 1];
                r.z = gr[3 * v + 2];
                return r;
            };
            const Dual S = sa_source<Dual>(m_rho(P, p, T), m_visc(P, T), nut, gvec(1), gvec(2), gvec(3), gvec(5), g.wd[c]);
            dSRC[c] = TD(g.vol[c]) * S.d;
        }
    }
    Dual q0[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) q0[v] = Dual(W[v * ld + c], dW[v * ld + c]);
    if constexpr (LIN) {
        static_assert(KNOWN, "the linearised limiter belongs to the known-state product");
        if (LIM != FCFD_LIMITER_NONE) {
            TD pd[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int j = lin.nb[v * ld + c];
                TD d = TD(0);
                if (j >= 0) d = TD(lin.a[v * ld + c]) * (dW[v * ld + j] - q0[v].d);
                d = d + (TD(lin.gv[(3 * v) * ld + c]) * gd[3 * v] + TD(lin.gv[(3 * v + 1) * ld + c]) * gd[3 * v + 1] +
                         TD(lin.gv[(3 * v + 2) * ld + c]) * gd[3 * v + 2]);
                pd[v] = d;
            }
            store_pairs<NV>(dPHI, ld, c, pd);
        }
    } else {
        Dual phi[NV];
        limiter_phi<LIM, NV>(g, P, TangentNb<TD>{W, dW, ld}, c, q0, gr, dref, phi);
        if (LIM != FCFD_LIMITER_NONE) {
            double pv[NV];
            TD pd[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { pv[v] = phi[v].v; pd[v] = phi[v].d; }
            if constexpr (!KNOWN) store_pairs<NV>(PHI, ld, c, pv);
            store_pairs<NV>(dPHI, ld, c, pd);
        }
    }
    // boundary ghosts take their parent's gradient
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) {
        const int h = h0 + min(k, nh - 1), fs = g.hf_face[h], gi = g.hf_nb[h];
        if (k < nh && fs >= 0 && gi >= g.N) {
            if constexpr (!KNOWN) store_pairs<3 * NV>(G, ld, gi, gv);
            store_pairs<3 * NV>(dG, ld, gi, gd);
        }
    }
}

// one wavefront per workgroup, as the primal face kernel (3% faster here)
template <int FLUX, int VISC, bool FULL, int NV, typename TD, bool PSEED = false>
__global__ void __launch_bounds__(64) face_flux_t_kernel(DevGrid g, Params P, DualLoader<NV, FULL, TD> L, TD* __restrict__ dFX,
                                                          int* __restrict__ status) {
    const int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    FaceCtx<DualT<TD>, NV> X;
    face_base<FLUX, VISC, NV, DualT<TD>, DualLoader<NV, FULL, TD>, PSEED>(g, P, L, f, X, status);
    const size_t Fs = g.F;
#pragma unroll
    for (int v = 0; v < NV; ++v) dFX[v * Fs + f] = (X.D[v] - X.C[v]).d;
}

// tangent of V_c S_c (SA source) added to the nu~ row; FULL carries the
// gradient tangents, FROZEN only the cell state's
template <bool FULL, typename TD>
__global__ void __launch_bounds__(256) sa_source_t_kernel(DevGrid g, Params P, const double* __restrict__ W, const TD* __restrict__ dW,
                                                          const double* __restrict__ G, const TD* __restrict__ dG,
                                                          TD* __restrict__ dR) {
    using Dual = DualT<TD>;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    Dual p(W[c], dW[c]), T(W[4 * ld + c], dW[4 * ld + c]), nut(W[5 * ld + c], dW[5 * ld + c]);
    auto gvec = [&](int v) {
        Vec<Dual> r;
        r.x = Dual(G[pidx(ld, 3 * v, c)], FULL ? dG[pidx(ld, 3 * v, c)] : TD(0));
        r.y = Dual(G[pidx(ld, 3 * v + 1, c)], FULL ? dG[pidx(ld, 3 * v + 1, c)] : TD(0));
        r.z = Dual(G[pidx(ld, 3 * v + 2, c)], FULL ? dG[pidx(ld, 3 * v + 2, c)] : TD(0));
        return r;
    };
    Dual S = sa_source<Dual>(m_rho(P, p, T), m_visc(P, T), nut, gvec(1), gvec(2), gvec(3), gvec(5), g.wd[c]);
    dR[5 * ld + c] += TD(g.vol[c]) * S.d;
}

// y = (P V/dt) x - dR  (the implicit operator applied matrix-free,
// time_terms ns_time_terms.cc:26-56,147-163 minus the tangent)
template <int NV, typename V>
__global__ void __launch_bounds__(256) newton_op_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ dt,
                                                        const V* __restrict__ x, const V* __restrict__ dR,
                                                        V* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double p = W[c], u = W[ld + c], v = W[2 * ld + c], w = W[3 * ld + c], T = W[4 * ld + c];
    double rho = m_rho(P, p, T);
    double pressure = p + P.Pref, temperature = T + P.Tref;
    double drho_dT = -1. * rho / temperature, drho_dp = rho / pressure;
    double H = P.Cp * temperature + 0.5 * (u * u + v * v + w * w);
    double xv[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) xv[k] = x[k * ld + c];
    double s = g.vol[c] / dt[c];
    double r0 = drho_dp * xv[0] + drho_dT * xv[4];
    double r1 = drho_dp * u * xv[0] + rho * xv[1] + drho_dT * u * xv[4];
    double r2 = drho_dp * v * xv[0] + rho * xv[2] + drho_dT * v * xv[4];
    double r3 = drho_dp * w * xv[0] + rho * xv[3] + drho_dT * w * xv[4];
    double r4 = (drho_dp * H - 1.) * xv[0] + rho * u * xv[1] + rho * v * xv[2] + rho * w * xv[3] + (drho_dT * H + rho * P.Cp) * xv[4];
    y[c] = V(s * r0 - dR[c]);
    y[ld + c] = V(s * r1 - dR[ld + c]);
    y[2 * ld + c] = V(s * r2 - dR[2 * ld + c]);
    y[3 * ld + c] = V(s * r3 - dR[3 * ld + c]);
    y[4 * ld + c] = V(s * r4 - dR[4 * ld + c]);
    if constexpr (NV == 6) {
        double nut = W[5 * ld + c];
        double r5 = drho_dp * nut * xv[0] + drho_dT * nut * xv[4] + rho * xv[5];
        y[5 * ld + c] = V(s * r5 - dR[5 * ld + c]);
    }
}

// y = (P V/dt)^T x - Jt  (the transposed implicit operator for the adjoint
// pseudo-time step; Jt = J^T x from the adjoint kernels).  The time-term block
// is the one newton_op_kernel applies, transposed.
template <int NV>
__global__ void __launch_bounds__(256) newton_op_t_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ dt,
                                                          const double* __restrict__ x, const double* __restrict__ Jt,
                                                          double* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double p = W[c], u = W[ld + c], v = W[2 * ld + c], w = W[3 * ld + c], T = W[4 * ld + c];
    double rho = m_rho(P, p, T);
    double pressure = p + P.Pref, temperature = T + P.Tref;
    double drho_dT = -1. * rho / temperature, drho_dp = rho / pressure;
    double H = P.Cp * temperature + 0.5 * (u * u + v * v + w * w);
    double xv[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) xv[k] = x[k * ld + c];
    double s = g.vol[c] / dt[c];
    double nut = 0., x5 = 0.;
    if constexpr (NV == 6) { nut = W[5 * ld + c]; x5 = xv[5]; }
    double c0 = drho_dp * xv[0] + drho_dp * u * xv[1] + drho_dp * v * xv[2] + drho_dp * w * xv[3] + (drho_dp * H - 1.) * xv[4] +
                drho_dp * nut * x5;
    double c1 = rho * xv[1] + rho * u * xv[4];
    double c2 = rho * xv[2] + rho * v * xv[4];
    double c3 = rho * xv[3] + rho * w * xv[4];
    double c4 = drho_dT * xv[0] + drho_dT * u * xv[1] + drho_dT * v * xv[2] + drho_dT * w * xv[3] + (drho_dT * H + rho * P.Cp) * xv[4] +
                drho_dT * nut * x5;
    y[c] = s * c0 - Jt[c];
    y[ld + c] = s * c1 - Jt[ld + c];
    y[2 * ld + c] = s * c2 - Jt[2 * ld + c];
    y[3 * ld + c] = s * c3 - Jt[3 * ld + c];
    y[4 * ld + c] = s * c4 - Jt[4 * ld + c];
    if constexpr (NV == 6) y[5 * ld + c] = s * (rho * x5) - Jt[5 * ld + c];
}

// ---------------------------------------------------------------------------
template <int FLUX, int VISC, bool FULL, int NV, typename TD>
static void launch_ft(const DevGrid& g, const Params& P, const DualLoader<NV, FULL, TD>& L, TD* dFX, int* status, hipStream_t s) {
    hipLaunchKernelGGL((face_flux_t_kernel<FLUX, VISC, FULL, NV, TD>), dim3((g.F + 63) / 64), dim3(64), 0, s, g, P, L, dFX, status);
}
template <int FLUX, bool FULL, int NV, typename TD>
static void launch_fv(const DevGrid& g, const Params& P, const DualLoader<NV, FULL, TD>& L, TD* dFX, int* status, hipStream_t s) {
    if constexpr (NV == 6) launch_ft<FLUX, 1, FULL, 6, TD>(g, P, L, dFX, status, s);   // SA is always viscous
    else if (P.kw) launch_ft<FLUX, kViscRans, FULL, 5, TD>(g, P, L, dFX, status, s);
    else if (P.viscous) launch_ft<FLUX, 1, FULL, 5, TD>(g, P, L, dFX, status, s);
    else launch_ft<FLUX, 0, FULL, 5, TD>(g, P, L, dFX, status, s);
}
template <bool FULL, int NV, typename TD>
static void launch_ff(const DevGrid& g, const Params& P, const DualLoader<NV, FULL, TD>& L, TD* dFX, int* status, hipStream_t s) {
    switch (P.flux) {
    case FCFD_FLUX_ROE: launch_fv<FCFD_FLUX_ROE, FULL, NV, TD>(g, P, L, dFX, status, s); break;
    case FCFD_FLUX_VAN_LEER: launch_fv<FCFD_FLUX_VAN_LEER, FULL, NV, TD>(g, P, L, dFX, status, s); break;
    case FCFD_FLUX_SD_SLAU: launch_fv<FCFD_FLUX_SD_SLAU, FULL, NV, TD>(g, P, L, dFX, status, s); break;
    case FCFD_FLUX_SW: launch_fv<FCFD_FLUX_SW, FULL, NV, TD>(g, P, L, dFX, status, s); break;
    default: launch_fv<FCFD_FLUX_AUSM_PLUS_UP, FULL, NV, TD>(g, P, L, dFX, status, s); break;
    }
}

template <int LIM, int NV, typename TD>
static void launch_glt(const DevGrid& g, const Params& P, double* W, TD* dW, double* G, TD* dG, double* PHI, TD* dPHI,
                       double* dref, bool known, TD* dsrc, hipStream_t s, const LimLin* lin = nullptr) {
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    const LimLin L = lin ? *lin : LimLin{nullptr, nullptr, nullptr};
    if constexpr (std::is_same<TD, float>::value && LIM != FCFD_LIMITER_NONE) {
        if (lin) {
            hipLaunchKernelGGL((grad_lim_t_kernel<LIM, NV, TD, true, true>), grid, blk, 0, s, g, P, W, dW, G, dG, PHI, dPHI, dref, dsrc, L);
            return;
        }
    }
    if (known) hipLaunchKernelGGL((grad_lim_t_kernel<LIM, NV, TD, true>), grid, blk, 0, s, g, P, W, dW, G, dG, PHI, dPHI, dref, dsrc, L);
    else hipLaunchKernelGGL((grad_lim_t_kernel<LIM, NV, TD, false>), grid, blk, 0, s, g, P, W, dW, G, dG, PHI, dPHI, dref, dsrc, L);
}

template <int NV, typename TD>
static void launch_jvp_t(const DevGrid& g, const Params& P, double* W, TD* dW, double* G, TD* dG, double* PHI, TD* dPHI,
                         double* dref, TD* dFX, TD* dR, bool full, int* status, hipStream_t s, bool known, TD* dsrc,
                         const LimLin* lin) {
    if (NV != 6 || !full) dsrc = nullptr;
    const int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    if (lim == FCFD_LIMITER_NONE || !full) lin = nullptr;
    const bool has_phi = lim != FCFD_LIMITER_NONE;
    const double phi_const = (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0;
    CellLoader<NV> cl{W, G, PHI, has_phi, phi_const, (size_t)g.ld};
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (full) {
        if (g.B > 0) hipLaunchKernelGGL((ghost_t_kernel<NV, TD>), dim3((g.B + 255) / 256), dim3(256), 0, s, g, P, W, dW, status);
        if constexpr (std::is_same<TD, double>::value) {   // (the FP32 tangent is refused on meshes with Green-Gauss cells)
            if (!known) launch_green_gauss(g, W, G, 0, s);
            launch_green_gauss(g, dW, dG, 1, s, &P.seed);
        }
        if (lin && !known) {   // the primal gradients, limiters and their linearisation at this W
            launch_grad_lim(g, P, W, G, PHI, dref, s);
            launch_lim_lin(g, P, W, G, dref, *lin, s);
            known = true;
        }
        if (lim == FCFD_LIMITER_VK) {
            if (!known) launch_vk_range(g, P, W, dref, s);   // primal reference range (constant in the tangent)
            launch_glt<FCFD_LIMITER_VK, NV, TD>(g, P, W, dW, G, dG, PHI, dPHI, dref, known, dsrc, s, lin);
        } else if (lim == FCFD_LIMITER_BJ) {
            launch_glt<FCFD_LIMITER_BJ, NV, TD>(g, P, W, dW, G, dG, PHI, dPHI, dref, known, dsrc, s, lin);
        } else if (lim == FCFD_LIMITER_MINMOD) {
            launch_glt<FCFD_LIMITER_MINMOD, NV, TD>(g, P, W, dW, G, dG, PHI, dPHI, dref, known, dsrc, s, lin);
        } else {
            launch_glt<FCFD_LIMITER_NONE, NV, TD>(g, P, W, dW, G, dG, PHI, dPHI, dref, known, dsrc, s);
        }
        if (!known) launch_wide_grad(g, G, g.wgrad, s);   // the primal face gradients of W (rewritten with G)
        if constexpr (std::is_same<TD, float>::value) launch_wide_grad(g, dG, g.wgrad_f, s);
        else launch_wide_grad(g, dG, g.wgrad_d, s);
        DualLoader<NV, true, TD> L{cl, dW, dG, dPHI, g.N};
        launch_ff<true, NV, TD>(g, P, L, dFX, status, s);
    } else {
        if (!known) {
            launch_ghosts(g, P, W, status, s);
            launch_grad_lim(g, P, W, G, PHI, dref, s);
        }
        DualLoader<NV, false, TD> L{cl, dW, nullptr, nullptr, g.N};
        launch_ff<false, NV, TD>(g, P, L, dFX, status, s);
    }
    if (dsrc) {   // the SA source tangent came from grad_lim_t_kernel
        launch_residual_sum(g, dsrc, dFX, dR, s);
        return;
    }
    launch_cell_sum(g, dFX, dR, s);
    if constexpr (NV == 6) {
        if (full) hipLaunchKernelGGL((sa_source_t_kernel<true, TD>), grid, blk, 0, s, g, P, W, dW, G, dG, dR);
        else hipLaunchKernelGGL((sa_source_t_kernel<false, TD>), grid, blk, 0, s, g, P, W, dW, G, dG, dR);
    }
}

void launch_jvp(const DevGrid& g, const Params& P, double* W, double* dW, double* G, double* dG, double* PHI, double* dPHI,
                double* dref, double* dFX, double* dR, bool full, int* status, hipStream_t s, bool primal_known, double* dsrc) {
    if (g.nv == 6) launch_jvp_t<6, double>(g, P, W, dW, G, dG, PHI, dPHI, dref, dFX, dR, full, status, s, primal_known, dsrc, nullptr);
    else launch_jvp_t<5, double>(g, P, W, dW, G, dG, PHI, dPHI, dref, dFX, dR, full, status, s, primal_known, dsrc, nullptr);
}

void launch_jvp(const DevGrid& g, const Params& P, double* W, float* dW, double* G, float* dG, double* PHI, float* dPHI,
                double* dref, float* dFX, float* dR, bool full, int* status, hipStream_t s, bool primal_known, float* dsrc,
                const LimLin* lin) {
    if (g.nv == 6) launch_jvp_t<6, float>(g, P, W, dW, G, dG, PHI, dPHI, dref, dFX, dR, full, status, s, primal_known, dsrc, lin);
    else launch_jvp_t<5, float>(g, P, W, dW, G, dG, PHI, dPHI, dref, dFX, dR, full, status, s, primal_known, dsrc, lin);
}

void launch_newton_op_t(const DevGrid& g, const Params& P, const double* W, const double* dt, const double* x, const double* Jt,
                        double* y, hipStream_t s) {
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL(newton_op_t_kernel<6>, grid, blk, 0, s, g, P, W, dt, x, Jt, y);
    else hipLaunchKernelGGL(newton_op_t_kernel<5>, grid, blk, 0, s, g, P, W, dt, x, Jt, y);
}

template <typename T>
static void launch_newton_op_v(const DevGrid& g, const Params& P, const double* W, const double* dt, const T* x, const T* dR, T* y,
                               hipStream_t s) {
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL((newton_op_kernel<6, T>), grid, blk, 0, s, g, P, W, dt, x, dR, y);
    else hipLaunchKernelGGL((newton_op_kernel<5, T>), grid, blk, 0, s, g, P, W, dt, x, dR, y);
}
void launch_newton_op(const DevGrid& g, const Params& P, const double* W, const double* dt, const double* x, const double* dR,
                      double* y, hipStream_t s) {
    launch_newton_op_v<double>(g, P, W, dt, x, dR, y, s);
}
void launch_newton_op(const DevGrid& g, const Params& P, const double* W, const double* dt, const float* x, const float* dR,
                      float* y, hipStream_t s) {
    launch_newton_op_v<float>(g, P, W, dt, x, dR, y, s);
}

}  // namespace fcfd

namespace fcfd {

// ---------------------------------------------------------------------------
// dual (pseudo) time stepping (ns_time_terms.cc:58-186)
// ---------------------------------------------------------------------------
// drho/dp of the pseudo-time block of each interior cell: cons2prim's rho/p,
// or with WS95 (preconditioner_ws95, :58-105) 1/Ur^2 - drho/dT / (rho c_p),
// Ur = min(max(|V|, 1e-5 a), a), raised to mu / (rho L) when mu > 1e-8 and to
// 1e-5 sqrt(dpmax / rho) (dpmax over the neighbour cells, ghosts included).
// With corr, R -= (P V/dt)(W - Wn) (the pseudo right-hand side, :117-146,
// 181-184); P is cons2prim at W.
template <int NV>
__global__ void __launch_bounds__(256) pseudo_rhs_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                         const double* __restrict__ Wn, const double* __restrict__ dt,
                                                         const int* __restrict__ nbp, const int* __restrict__ nbi, int ws95,
                                                         int corr, double* __restrict__ R, double* __restrict__ drdp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    const double p = W[c], u = W[ld + c], v = W[2 * ld + c], w = W[3 * ld + c], T = W[4 * ld + c];
    const double rho = m_rho(P, p, T);
    const double pressure = p + P.Pref, temperature = T + P.Tref;
    const double drho_dT = -1. * rho / temperature, drho_dp = rho / pressure;
    double d = drho_dp;
    if (ws95) {
        const double a = m_a(P, p, T);
        double Ur = smax(sqrt(u * u + v * v + w * w), 1.e-5 * a);
        Ur = smin(Ur, a);
        const double mu = m_visc(P, T);
        if (mu > 1.e-8) Ur = smax(Ur, mu / (rho * g.lscale[c]));
        double dpmax = 0.;
        for (int k = nbp[c]; k < nbp[c + 1]; ++k) dpmax = smax(dpmax, fabs(p - W[nbi[k]]));
        Ur = smax(Ur, 1.e-5 * sqrt(dpmax / rho));
        d = 1. / (Ur * Ur) - drho_dT / (rho * P.Cp);
    }
    drdp[c] = d;
    if (!corr) return;
    const double H = P.Cp * temperature + 0.5 * (u * u + v * v + w * w);
    const double s = g.vol[c] / dt[c];
    double x[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) x[k] = W[k * ld + c] - Wn[k * ld + c];
    const double row[6][6] = {{drho_dp, 0., 0., 0., drho_dT, 0.},
                              {drho_dp * u, rho, 0., 0., drho_dT * u, 0.},
                              {drho_dp * v, 0., rho, 0., drho_dT * v, 0.},
                              {drho_dp * w, 0., 0., rho, drho_dT * w, 0.},
                              {drho_dp * H - 1., rho * u, rho * v, rho * w, drho_dT * H + rho * P.Cp, 0.},
                              {NV == 6 ? drho_dp * W[5 * ld + c] : 0., 0., 0., 0., NV == 6 ? drho_dT * W[5 * ld + c] : 0., rho}};
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double corr_i = 0.;
#pragma unroll
        for (int j = 0; j < NV; ++j) corr_i += (row[i][j] * s) * x[j];
        R[i * ld + c] -= corr_i;
    }
}

// the pseudo-time block P_ps V/dtau (P_ps: cons2prim with drho/dp = drdp[c])
// of each interior cell, as a dense nv x nv array (row-major)
template <int NV>
__device__ __forceinline__ void pseudo_block(const DevGrid& g, const Params& P, const double* __restrict__ W, double dtau, double d,
                                             int c, double (&B)[NV][NV]) {
    const size_t ld = g.ld;
    const double p = W[c], u = W[ld + c], v = W[2 * ld + c], w = W[3 * ld + c], T = W[4 * ld + c];
    const double rho = m_rho(P, p, T);
    const double temperature = T + P.Tref;
    const double drho_dT = -1. * rho / temperature;
    const double H = P.Cp * temperature + 0.5 * (u * u + v * v + w * w);
    const double s = g.vol[c] / dtau;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) B[i][j] = 0.;
    B[0][0] = d * s; B[0][4] = drho_dT * s;
    B[1][0] = d * u * s; B[1][1] = rho * s; B[1][4] = drho_dT * u * s;
    B[2][0] = d * v * s; B[2][2] = rho * s; B[2][4] = drho_dT * v * s;
    B[3][0] = d * w * s; B[3][3] = rho * s; B[3][4] = drho_dT * w * s;
    B[4][0] = (d * H - 1.) * s; B[4][1] = rho * u * s; B[4][2] = rho * v * s; B[4][3] = rho * w * s;
    B[4][4] = (drho_dT * H + rho * P.Cp) * s;
    if constexpr (NV == 6) {
        const double nut = W[5 * ld + c];
        B[5][0] = d * nut * s; B[5][4] = drho_dT * nut * s; B[5][5] = rho * s;
    }
}

// D += P_ps V/dtau (diagonal blocks in row pairs, entry k = nv row + col)
template <int NV>
__global__ void __launch_bounds__(256) pseudo_diag_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                          const double* __restrict__ dtau, const double* __restrict__ drdp,
                                                          double* __restrict__ D) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    double B[NV][NV];
    pseudo_block<NV>(g, P, W, dtau[c], drdp[c], c, B);
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) D[pidx(g.ld, NV * i + j, c)] += B[i][j];
}

// y += P_ps V/dtau x (the matrix-free operator's pseudo-time term)
template <int NV, typename T>
__global__ void __launch_bounds__(256) pseudo_op_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ dtau,
                                                        const double* __restrict__ drdp, const T* __restrict__ x,
                                                        T* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double B[NV][NV];
    pseudo_block<NV>(g, P, W, dtau[c], drdp[c], c, B);
    double xv[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) xv[j] = x[j * ld + c];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = 0.;
#pragma unroll
        for (int j = 0; j < NV; ++j) s += B[i][j] * xv[j];
        y[i * ld + c] = T(y[i * ld + c] + s);
    }
}

void launch_pseudo_rhs(const DevGrid& g, const Params& P, const double* W, const double* Wn, const double* dt, const int* nbp,
                       const int* nbi, int ws95, int corr, double* R, double* drdp, hipStream_t s) {
    const dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL(pseudo_rhs_kernel<6>, grid, blk, 0, s, g, P, W, Wn, dt, nbp, nbi, ws95, corr, R, drdp);
    else hipLaunchKernelGGL(pseudo_rhs_kernel<5>, grid, blk, 0, s, g, P, W, Wn, dt, nbp, nbi, ws95, corr, R, drdp);
}

void launch_pseudo_diag(const DevGrid& g, const Params& P, const double* W, const double* dtau, const double* drdp, double* D,
                        hipStream_t s) {
    const dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL(pseudo_diag_kernel<6>, grid, blk, 0, s, g, P, W, dtau, drdp, D);
    else hipLaunchKernelGGL(pseudo_diag_kernel<5>, grid, blk, 0, s, g, P, W, dtau, drdp, D);
}

template <typename T>
static void launch_pseudo_op_v(const DevGrid& g, const Params& P, const double* W, const double* dtau, const double* drdp, const T* x,
                               T* y, hipStream_t s) {
    const dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL((pseudo_op_kernel<6, T>), grid, blk, 0, s, g, P, W, dtau, drdp, x, y);
    else hipLaunchKernelGGL((pseudo_op_kernel<5, T>), grid, blk, 0, s, g, P, W, dtau, drdp, x, y);
}
void launch_pseudo_op(const DevGrid& g, const Params& P, const double* W, const double* dtau, const double* drdp, const double* x,
                      double* y, hipStream_t s) {
    launch_pseudo_op_v<double>(g, P, W, dtau, drdp, x, y, s);
}
void launch_pseudo_op(const DevGrid& g, const Params& P, const double* W, const double* dtau, const double* drdp, const float* x,
                      float* y, hipStream_t s) {
    launch_pseudo_op_v<float>(g, P, W, dtau, drdp, x, y, s);
}


// ---------------------------------------------------------------------------
// the parameter tangent: dR and dFX at fixed W in the direction P.seed of the
// values of one BC (every path from a BC value: ghost centres, then
// gradients and limiters, the boundary faces' right states, a FIXED_Q heat
// flux, the SA inlet nu~).  dW must be 0 on interior cells.
// ---------------------------------------------------------------------------
template <int FLUX, int NV>
static void launch_fp(const DevGrid& g, const Params& P, const DualLoader<NV, true, double>& L, double* dFX, int* status, hipStream_t s) {
    const dim3 grid((g.F + 63) / 64), blk(64);
    if constexpr (NV == 6)
        hipLaunchKernelGGL((face_flux_t_kernel<FLUX, 1, true, 6, double, true>), grid, blk, 0, s, g, P, L, dFX, status);
    else if (P.kw)
        hipLaunchKernelGGL((face_flux_t_kernel<FLUX, kViscRans, true, 5, double, true>), grid, blk, 0, s, g, P, L, dFX, status);
    else if (P.viscous)
        hipLaunchKernelGGL((face_flux_t_kernel<FLUX, 1, true, 5, double, true>), grid, blk, 0, s, g, P, L, dFX, status);
    else
        hipLaunchKernelGGL((face_flux_t_kernel<FLUX, 0, true, 5, double, true>), grid, blk, 0, s, g, P, L, dFX, status);
}

template <int NV>
static void launch_param_t(const DevGrid& g, const Params& P, double* W, double* dW, double* G, double* dG, double* PHI, double* dPHI,
                           double* dref, double* dFX, double* dR, int* status, hipStream_t s) {
    const int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    const bool has_phi = lim != FCFD_LIMITER_NONE;
    const double phi_const = (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0;
    CellLoader<NV> cl{W, G, PHI, has_phi, phi_const, (size_t)g.ld};
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.B > 0) hipLaunchKernelGGL((ghost_t_kernel<NV, double, true>), dim3((g.B + 255) / 256), dim3(256), 0, s, g, P, W, dW, status);
    launch_green_gauss(g, W, G, 0, s);
    launch_green_gauss(g, dW, dG, 1, s, &P.seed);
    if (lim == FCFD_LIMITER_VK) {
        launch_vk_range(g, P, W, dref, s);
        launch_glt<FCFD_LIMITER_VK, NV, double>(g, P, W, dW, G, dG, PHI, dPHI, dref, false, nullptr, s);
    } else if (lim == FCFD_LIMITER_BJ) {
        launch_glt<FCFD_LIMITER_BJ, NV, double>(g, P, W, dW, G, dG, PHI, dPHI, dref, false, nullptr, s);
    } else if (lim == FCFD_LIMITER_MINMOD) {
        launch_glt<FCFD_LIMITER_MINMOD, NV, double>(g, P, W, dW, G, dG, PHI, dPHI, dref, false, nullptr, s);
    } else {
        launch_glt<FCFD_LIMITER_NONE, NV, double>(g, P, W, dW, G, dG, PHI, dPHI, dref, false, nullptr, s);
    }
    launch_wide_grad(g, G, g.wgrad, s);
    launch_wide_grad(g, dG, g.wgrad_d, s);
    DualLoader<NV, true, double> L{cl, dW, dG, dPHI, g.N};
    switch (P.flux) {
    case FCFD_FLUX_ROE: launch_fp<FCFD_FLUX_ROE, NV>(g, P, L, dFX, status, s); break;
    case FCFD_FLUX_VAN_LEER: launch_fp<FCFD_FLUX_VAN_LEER, NV>(g, P, L, dFX, status, s); break;
    case FCFD_FLUX_SD_SLAU: launch_fp<FCFD_FLUX_SD_SLAU, NV>(g, P, L, dFX, status, s); break;
    case FCFD_FLUX_SW: launch_fp<FCFD_FLUX_SW, NV>(g, P, L, dFX, status, s); break;
    default: launch_fp<FCFD_FLUX_AUSM_PLUS_UP, NV>(g, P, L, dFX, status, s); break;
    }
    launch_cell_sum(g, dFX, dR, s);
    if constexpr (NV == 6) hipLaunchKernelGGL((sa_source_t_kernel<true, double>), grid, blk, 0, s, g, P, W, dW, G, dG, dR);
}

void launch_param_tangent(const DevGrid& g, const Params& P, double* W, double* dW, double* G, double* dG, double* PHI, double* dPHI,
                          double* dref, double* dFX, double* dR, int* status, hipStream_t s) {
    if (g.nv == 6) launch_param_t<6>(g, P, W, dW, G, dG, PHI, dPHI, dref, dFX, dR, status, s);
    else launch_param_t<5>(g, P, W, dW, G, dG, PHI, dPHI, dref, dFX, dR, status, s);
}

}  // namespace fcfd
