// Residual kernels for gfx950 (FP64).
//
// Launches per residual evaluation, all deterministic and atomic-free (every
// kernel is templated on the number of variables per cell, NV = 5, or 6 with
// the SA model):
//   1. ghost_kernel      one thread per boundary face: boundary ghost state
//   2. grad_lim_kernel   one thread per cell: gradient-map gather (sliced
//                        ELL, 64-wide slices matching the wavefront) fused
//                        with the limiter (and, with SA, the cell's V S
//                        source); also copies gradients to the ghosts.  Only
//                        when some cells use Green-Gauss gradients (no map),
//                        gg_kernel computes theirs just before it.
//   3. face_flux_kernel  one thread per face, one wavefront per workgroup:
//                        reconstruction, boundary condition, convective and
//                        diffusive flux; D - C per face into FX [nv][F]
//   4. cell_sum_kernel   one thread per cell: the cell's face contributions
//                        in reference face order (the reference's
//                        VecSetValues(ADD) sequence), so R rounds as the
//                        reference's does
// The gather kernels are bound by chains of dependent loads (an index, then
// what it points at), so each issues its loads in as few dependent rounds as
// it can: all of a cell's half-face entries at once (at most kMaxCellFaces,
// indices past the count clamped to a valid slot and the arithmetic masked),
// a face's packed records together before the cell loads they index.
#include <hip/hip_runtime.h>

#include "fcfd_device.h"
#include "fcfd_face.h"

namespace fcfd {

// --------------------------------------------------------------------------
// 1. boundary ghosts: apply_bcs with the parent's cell-centre state as left and
//    right pre-set to a copy of left; ghost = right.{p,V,T}_center
// --------------------------------------------------------------------------
template <int NV>
__global__ void __launch_bounds__(256) ghost_kernel(DevGrid g, Params P, double* __restrict__ W, int* __restrict__ status) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
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
    if constexpr (NV == 6) L.nut = L.nut_center = W[5 * ld + c];
    else L.nut = L.nut_center = 0.0;
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
    if constexpr (NV == 6) W[5 * ld + gi] = R.nut_center;
}

// --------------------------------------------------------------------------
// 2. gradients (calc_cell_grads, ns.cc:301-336 + variable.h:193-216) fused
//    with the limiter (ns_limiters.cc:26-248)
// --------------------------------------------------------------------------
// Green-Gauss gradient of each cell without a gradient map
// (Variable::cell_gradient, variable.h:193-216): over the cell's faces in
// reference order, the face value times n A / V, negated where the cell is
// not the face's parent.  The face value is the BC value where the variable
// is fixed on the face's BC, else the face average w_p q_p + w_n q_n over the
// parent and the neighbour or ghost (Variable::face_calculate,
// variable.h:135-148).  Runs after the ghost kernel and before
// grad_lim_kernel, which takes these gradients for its limiters and ghost
// copies.  The sum is linear in the cell values, so with tangent = 1 the
// same kernel on a tangent dW gives the tangent gradient (a fixed BC value
// contributes 0).
template <int NV>
__global__ void __launch_bounds__(256) gg_kernel(DevGrid g, const double* __restrict__ W, double* __restrict__ G, int tangent, BcSeed seed) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N || !g.ggc[c]) return;
    const size_t ld = g.ld;
    const double vol = g.vol[c];
    double gr[3 * NV];
#pragma unroll
    for (int k = 0; k < 3 * NV; ++k) gr[k] = 0.0;
    for (int h = g.hf_ptr[c]; h < g.hf_ptr[c + 1]; ++h) {
        const int fs = g.hf_face[h];
        const int f = fs >= 0 ? fs : ~fs;
        const double A = g.farea[f];
        double a0 = g.fnx[f] * A / vol, a1 = g.fny[f] * A / vol, a2 = g.fnz[f] * A / vol;
        if (fs < 0) { a0 = a0 * -1.; a1 = a1 * -1.; a2 = a2 * -1.; }
        const int4 fc = g.fcell[f];
        const double2 w = g.ggw[f];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double fv;
            bool fixed = false;
            if (fc.z >= 0) {
                const BcDev& bc = g.bc[fc.z];
                if (v == 0 && bc.p_fixed) { fv = bc.p; fixed = true; }
                if (v == 4 && bc.T_fixed) { fv = bc.T; fixed = true; }
                if (v >= 1 && v <= 3 && bc.V_fixed) { fv = bc.V[v - 1]; fixed = true; }
                // the tangent of a fixed face value: 0, or the seeded BC value's direction
                if (fixed && tangent)
                    fv = fc.z != seed.index ? 0.0 : v == 0 ? seed.p : v == 4 ? seed.T : seed.V[v - 1];
            }
            if (!fixed) {
                double t = 0.;
                t += w.x * W[v * ld + fc.x];
                t += w.y * W[v * ld + fc.y];
                fv = t;
            }
            gr[3 * v] += fv * a0;
            gr[3 * v + 1] += fv * a1;
            gr[3 * v + 2] += fv * a2;
        }
    }
    store_pairs<3 * NV>(G, ld, c, gr);
}

// (3 waves per SIMD: 168 VGPRs with a 12-byte spill; the compiler's own
// choice, 174 VGPRs at 2 waves, is 6% slower)
template <int LIM, int NV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) grad_lim_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                       double* __restrict__ G, double* __restrict__ PHI,
                                                       const double* __restrict__ dref, double* __restrict__ SRC) {
    // The loads are issued in rounds (all map entries of a chunk, then all
    // their states; all half-faces, then all neighbour states and face
    // centroids) rather than entry by entry: the kernel is bound by chains of
    // dependent loads.  The arithmetic is the reference's, in its order.
    constexpr int kChunk = 6;   // gradient-map entries per round
    int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double q0[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) q0[v] = W[v * ld + c];
    const int h0 = g.hf_ptr[c], nh = g.hf_ptr[c + 1] - h0;
    double gr[3 * NV];
#pragma unroll
    for (int k = 0; k < 3 * NV; ++k) gr[k] = 0.0;
    {   // entry k of cell `lane` of slice s at e0 + 64 k + lane (K entries each, zero-weight pads included)
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
            double q[kChunk][NV];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int j = (int)__double_as_longlong(b[u].y);
#pragma unroll
                for (int v = 0; v < NV; ++v) q[u][v] = W[v * ld + j];
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                if (k0 + u < K) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        gr[3 * v] += a[u].x * q[u][v];
                        gr[3 * v + 1] += a[u].y * q[u][v];
                        gr[3 * v + 2] += b[u].x * q[u][v];
                    }
                }
            }
        }
    }
    if (g.ggc && g.ggc[c]) load_pairs<3 * NV>(G, ld, c, gr);   // a Green-Gauss cell: gg_kernel's gradient
    store_pairs<3 * NV>(G, ld, c, gr);
    if constexpr (NV == 6) {
        // V_c S_c of the SA source (ns_sources.cc:25-30 has none; build-only model)
        if (SRC) {
            auto gv = [&](int v) { return mk<double>(gr[3 * v], gr[3 * v + 1], gr[3 * v + 2]); };
            SRC[c] = g.vol[c] * sa_source<double>(m_rho(P, q0[0], q0[4]), m_visc(P, q0[4]), q0[5], gv(1), gv(2), gv(3), gv(5),
                                                  g.wd[c]);
        }
    }
    // the cell's half-faces (at most kMaxCellFaces; slots past nh repeat the
    // last one and are masked)
    int fs[kMaxCellFaces], nb[kMaxCellFaces];
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) {
        const int h = h0 + min(k, nh - 1);
        fs[k] = g.hf_face[h];
        nb[k] = g.hf_nb[h];
    }
    double phi[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) phi[v] = 1.;
    if (LIM != FCFD_LIMITER_NONE) {
        const Vec<double> xc = centroid(g, c);
        // the neighbour states (min/max, or minmod's differences), then the
        // face centroids (so both sets are never live at once)
        if (LIM == FCFD_LIMITER_MINMOD) {
            double qn[kMaxCellFaces][NV];
#pragma unroll
            for (int k = 0; k < kMaxCellFaces; ++k) {
#pragma unroll
                for (int v = 0; v < NV; ++v) qn[k][v] = W[v * ld + nb[k]];
            }
            Vec<double> c2f[kMaxCellFaces];
#pragma unroll
            for (int k = 0; k < kMaxCellFaces; ++k) c2f[k] = vsub(face_centroid(g, fs[k] >= 0 ? fs[k] : ~fs[k]), xc);
            const double small = 1.e-12;
#pragma unroll
            for (int k = 0; k < kMaxCellFaces; ++k) {
                if (k >= nh) break;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    double dU = qn[k][v] - q0[v];
                    double dUg = gr[3 * v] * c2f[k].x + gr[3 * v + 1] * c2f[k].y + gr[3 * v + 2] * c2f[k].z;
                    double mm = (dU * dUg < 0.) ? 0. : (dU > 0.) ? smin(dU, dUg) : smax(dU, dUg);
                    double sU = (double)((mm > 0) - (mm < 0)), sG = (double)((dUg > 0) - (dUg < 0));
                    phi[v] = smin(phi[v], (mm + sU * small) / (dUg + sG * small));
                }
            }
        } else if (LIM == FCFD_LIMITER_BJ || LIM == FCFD_LIMITER_VK) {
            double umax[NV], umin[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) umax[v] = umin[v] = q0[v];
            // two rounds of half the faces (max and min are exact: the
            // grouping does not change them)
            constexpr int kHalf = kMaxCellFaces / 2;
#pragma unroll
            for (int k0 = 0; k0 < kMaxCellFaces; k0 += kHalf) {
                double qn[kHalf][NV];
#pragma unroll
                for (int u = 0; u < kHalf; ++u) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) qn[u][v] = W[v * ld + nb[k0 + u]];
                }
#pragma unroll
                for (int u = 0; u < kHalf; ++u) {
                    if (k0 + u < nh) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            umax[v] = smax(umax[v], qn[u][v]);
                            umin[v] = smin(umin[v], qn[u][v]);
                        }
                    }
                }
            }
            Vec<double> c2f[kMaxCellFaces];
#pragma unroll
            for (int k = 0; k < kMaxCellFaces; ++k) c2f[k] = vsub(face_centroid(g, fs[k] >= 0 ? fs[k] : ~fs[k]), xc);
#pragma unroll
            for (int k = 0; k < kMaxCellFaces; ++k) {
                if (k >= nh) break;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    double dM = gr[3 * v] * c2f[k].x + gr[3 * v + 1] * c2f[k].y + gr[3 * v + 2] * c2f[k].z;
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
        store_pairs<NV>(PHI, ld, c, phi);
    }
    // boundary ghosts take their parent's gradient (ns.cc:320-333)
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) {
        if (k < nh && fs[k] >= 0 && nb[k] >= g.N) store_pairs<3 * NV>(G, ld, nb[k], gr);
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
// slots: 0..2 max p,u,v   3: max p excluding `last`   4..6: min p,u,v   7: min
// p excluding `last`; with the SA model 8: max nu~, 9: min nu~ (a plain range:
// the build-only variable has no reference slot)
template <int NV>
__global__ void __launch_bounds__(256) minmax_partial_kernel(int N, size_t ld, int first, int last, const double* __restrict__ W,
                                                             double* __restrict__ part) {
    constexpr int K = NV == 6 ? 10 : 8;
    __shared__ double sh[K][256];
    int t = threadIdx.x;
    double r[K];
    r[0] = r[4] = W[first]; r[1] = r[5] = W[ld + first]; r[2] = r[6] = W[2 * ld + first];
    r[3] = r[7] = W[4 * ld + first];   // T_0 seeds the p-tracking slots
    if constexpr (NV == 6) r[8] = r[9] = W[5 * ld + first];
    for (int c = blockIdx.x * 256 + t; c < N; c += kMMBlocks * 256) {
        double p = W[c], u = W[ld + c], v = W[2 * ld + c];
        r[0] = smax(r[0], p); r[1] = smax(r[1], u); r[2] = smax(r[2], v);
        r[4] = smin(r[4], p); r[5] = smin(r[5], u); r[6] = smin(r[6], v);
        if (c != last) { r[3] = smax(r[3], p); r[7] = smin(r[7], p); }
        if constexpr (NV == 6) { double n = W[5 * ld + c]; r[8] = smax(r[8], n); r[9] = smin(r[9], n); }
    }
    for (int k = 0; k < K; ++k) sh[k][t] = r[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < K; ++k) sh[k][t] = (k < 4 || k == 8) ? smax(sh[k][t], sh[k][t + s]) : smin(sh[k][t], sh[k][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) for (int k = 0; k < K; ++k) part[k * kMMBlocks + blockIdx.x] = sh[k][0];
}

template <int NV>
__global__ void __launch_bounds__(256) minmax_final_kernel(size_t ld, int first, int last, const double* __restrict__ W,
                                                           const double* __restrict__ part, double* __restrict__ dref) {
    constexpr int K = NV == 6 ? 10 : 8;
    __shared__ double sh[K][256];
    int t = threadIdx.x;
    double r[K];
    for (int k = 0; k < K; ++k) r[k] = part[k * kMMBlocks + t];
    for (int i = t + 256; i < kMMBlocks; i += 256)
        for (int k = 0; k < K; ++k)
            r[k] = (k < 4 || k == 8) ? smax(r[k], part[k * kMMBlocks + i]) : smin(r[k], part[k * kMMBlocks + i]);
    for (int k = 0; k < K; ++k) sh[k][t] = r[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < K; ++k) sh[k][t] = (k < 4 || k == 8) ? smax(sh[k][t], sh[k][t + s]) : smin(sh[k][t], sh[k][t + s]);
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
        if constexpr (NV == 6) dref[5] = fabs(sh[8][0] - sh[9][0]);
    }
}

// --------------------------------------------------------------------------
// 3. face fluxes (ns_assemble_linear_system.cc:75-130): D - C per face
// --------------------------------------------------------------------------
// one wavefront per workgroup: at 2 waves per SIMD the CU then fills with
// single waves (5% faster than 4-wave workgroups)
template <int FLUX, int VISC, int NV>
__global__ void __launch_bounds__(64) face_flux_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                        const double* __restrict__ G, const double* __restrict__ PHI,
                                                        int has_phi, double phi_const, double* __restrict__ FX,
                                                        int* __restrict__ status) {
    const int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    FaceCtx<double, NV> X;
    CellLoader<NV> ld_{W, G, PHI, has_phi != 0, phi_const, (size_t)g.ld};
    face_base<FLUX, VISC, NV>(g, P, ld_, f, X, status);
    const size_t Fs = g.F;
#pragma unroll
    for (int v = 0; v < NV; ++v) FX[v * Fs + f] = X.D[v] - X.C[v];
}

// Surface quantities of the face loop (ns_assemble_linear_system.cc:122-127):
// mdot = (C - D)[0] / A on every face; on boundary faces qdot = (C - D)[4] / A
// (the given value where qdot is fixed on the BC, which the loop does not
// overwrite) and tau = -D[1..3] / A.  out: [5][F] rows mdot, qdot, tau_x,
// tau_y, tau_z (interior faces: qdot and tau rows 0).
template <int FLUX, int VISC, int NV>
__global__ void __launch_bounds__(64) surface_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ G,
                                                     const double* __restrict__ PHI, int has_phi, double phi_const,
                                                     double* __restrict__ out, int* __restrict__ status) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    FaceCtx<double, NV> X;
    CellLoader<NV> ld_{W, G, PHI, has_phi != 0, phi_const, (size_t)g.ld};
    face_base<FLUX, VISC, NV>(g, P, ld_, f, X, status);
    const size_t Fs = g.F;
    const double A = X.F.area;
    out[f] = (X.C[0] - X.D[0]) / A;
    double q = 0., tx = 0., ty = 0., tz = 0.;
    if (X.F.bc >= 0) {
        const BcDev& bc = g.bc[X.F.bc];
        q = bc.qdot_fixed ? bc.qdot : (X.C[4] - X.D[4]) / A;
        tx = -X.D[1] / A; ty = -X.D[2] / A; tz = -X.D[3] / A;
    }
    out[Fs + f] = q;
    out[2 * Fs + f] = tx;
    out[3 * Fs + f] = ty;
    out[4 * Fs + f] = tz;
}

// R_c = sum over the faces of c in reference face order of (D-C)+0 (parent)
// or -1*(D-C)+0 (neighbour) -- the VecSetValues(ADD) sequence of
// ns_assemble_linear_system.cc:131-140 starting from VecSet(rhs, 0) -- then,
// with SRC (SA model), + V_c S_c (src[c], written by grad_lim_kernel, or its
// tangent by grad_lim_t_kernel) in the nu~ row
template <int NV, bool SRC, typename T>
__global__ void __launch_bounds__(256) cell_sum_kernel(DevGrid g, const T* __restrict__ src, const T* __restrict__ FX,
                                                       T* __restrict__ Rout) {
    // all half-face entries in one round, then all their flux rows in one
    // round (entries past the cell's count repeat its last one and are
    // masked), then the sum in order
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld, Fs = g.F;
    const int h0 = g.hf_ptr[c], nh = g.hf_ptr[c + 1] - h0;
    int fs[kMaxCellFaces];
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) fs[k] = g.hf_face[h0 + min(k, nh - 1)];
    T x[kMaxCellFaces][NV];
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) {
        const int f = fs[k] >= 0 ? fs[k] : ~fs[k];
#pragma unroll
        for (int v = 0; v < NV; ++v) x[k][v] = FX[v * Fs + f];
    }
    T s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = T(0);
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) {
        if (k < nh) {
#pragma unroll
            for (int v = 0; v < NV; ++v) s[v] += (fs[k] >= 0 ? x[k][v] : T(-1) * x[k][v]) + T(0);
        }
    }
    if constexpr (SRC && NV == 6) s[5] += src[c];
#pragma unroll
    for (int v = 0; v < NV; ++v) Rout[v * ld + c] = s[v];
}

// --------------------------------------------------------------------------
// the convective flux alone, over a batch of faces (the reference's leaf
// functions roe_flux, AUSMplusUP_flux, vanLeer_flux, SD_SLAU_flux,
// Stegger_Warming_flux with convective_face_flux's rotation and area,
// ns_convective_face_flux.cc:25-59): each side's state from p, u, v, w, T
// as the face loop finishes it (rho by the EOS, a, H, Vn in the face frame
// n, t1, t2 = norm(n x t1)); flux [5][n], weightL [n]
// --------------------------------------------------------------------------
template <int FLUX>
__global__ void __launch_bounds__(256) flux_leaf_kernel(Params P, int n, const double* __restrict__ WL, const double* __restrict__ WR,
                                                        const double* __restrict__ nrm, const double* __restrict__ tan1,
                                                        const double* __restrict__ area, const double* __restrict__ diss,
                                                        const double* __restrict__ wdist, double* __restrict__ out,
                                                        double* __restrict__ wl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t m = (size_t)n;
    FaceGeo F;
    F.n = mk<double>(nrm[i], nrm[m + i], nrm[2 * m + i]);
    F.t1 = mk<double>(tan1[i], tan1[m + i], tan1[2 * m + i]);
    F.t2 = vnorm(vcross(F.n, F.t1));
    F.l2r = mk<double>(0., 0., 0.);
    F.area = area[i];
    F.bc = -1;
    CState<double> S[2];
    const double* W[2] = {WL, WR};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        S[k].p = S[k].p_center = W[k][i];
        S[k].V = S[k].V_center = mk<double>(W[k][m + i], W[k][2 * m + i], W[k][3 * m + i]);
        S[k].T = S[k].T_center = W[k][4 * m + i];
        S[k].nut = S[k].nut_center = 0.;
        S[k].rho = m_rho(P, S[k].p, S[k].T);
        finish_state(P, S[k], F);
    }
    double fl[5];
    convective_flux<FLUX, 5, double>(P, S[0], S[1], F, diss ? diss[i] : 1., wdist ? wdist[i] : 1.e20, fl);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v * m + i] = fl[v];
    if (wl) wl[i] = weight_left<FLUX>(S[0], S[1], fl[0]);
}

void launch_flux_leaf(const Params& P, int flux, int n, const double* WL, const double* WR, const double* nrm, const double* tan1,
                      const double* area, const double* diss, const double* wdist, double* out, double* wl, hipStream_t s) {
    const dim3 grid((n + 255) / 256), blk(256);
    switch (flux) {
    case FCFD_FLUX_ROE:
        hipLaunchKernelGGL(flux_leaf_kernel<FCFD_FLUX_ROE>, grid, blk, 0, s, P, n, WL, WR, nrm, tan1, area, diss, wdist, out, wl);
        break;
    case FCFD_FLUX_VAN_LEER:
        hipLaunchKernelGGL(flux_leaf_kernel<FCFD_FLUX_VAN_LEER>, grid, blk, 0, s, P, n, WL, WR, nrm, tan1, area, diss, wdist, out, wl);
        break;
    case FCFD_FLUX_SD_SLAU:
        hipLaunchKernelGGL(flux_leaf_kernel<FCFD_FLUX_SD_SLAU>, grid, blk, 0, s, P, n, WL, WR, nrm, tan1, area, diss, wdist, out, wl);
        break;
    case FCFD_FLUX_SW:
        hipLaunchKernelGGL(flux_leaf_kernel<FCFD_FLUX_SW>, grid, blk, 0, s, P, n, WL, WR, nrm, tan1, area, diss, wdist, out, wl);
        break;
    default:
        hipLaunchKernelGGL(flux_leaf_kernel<FCFD_FLUX_AUSM_PLUS_UP>, grid, blk, 0, s, P, n, WL, WR, nrm, tan1, area, diss, wdist, out, wl);
        break;
    }
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

// closest wall distance of every cell (set_bcs.cc:178-210 brute force): the
// minimum over the no-slip wall face centroids of |x_c - x_f|, exact and so
// independent of the order of the search; wall centroids staged in LDS tiles
__global__ void __launch_bounds__(256) wall_distance_kernel(int T, const double* __restrict__ ccx, const double* __restrict__ ccy,
                                                            const double* __restrict__ ccz, int nw, const double* __restrict__ wx,
                                                            const double* __restrict__ wy, const double* __restrict__ wz,
                                                            double* __restrict__ wd) {
    __shared__ double tx[256], ty[256], tz[256];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = c < T;
    Vec<double> x = active ? mk<double>(ccx[c], ccy[c], ccz[c]) : mk<double>(0., 0., 0.);
    double best = 1.e20;
    for (int t0 = 0; t0 < nw; t0 += 256) {
        int k = t0 + threadIdx.x;
        if (k < nw) { tx[threadIdx.x] = wx[k]; ty[threadIdx.x] = wy[k]; tz[threadIdx.x] = wz[k]; }
        __syncthreads();
        const int m = min(256, nw - t0);
        for (int q = 0; q < m; ++q) {
            double d = vmag(vsub(x, mk<double>(tx[q], ty[q], tz[q])));
            if (best > d) best = d;
        }
        __syncthreads();
    }
    if (active) wd[c] = best;
}

// closest no-slip wall face (centroid distance) of every non-wall face and the
// Steger-Warming dissipation factor 1 - |n_wall . n_f| of the first closest
// wall face in reference face order (set_bcs.cc:211-249); wall faces get 0, 0
__global__ void __launch_bounds__(256) face_wall_kernel(int F, const double* __restrict__ fcx, const double* __restrict__ fcy,
                                                        const double* __restrict__ fcz, const double* __restrict__ fnx,
                                                        const double* __restrict__ fny, const double* __restrict__ fnz,
                                                        const int* __restrict__ is_wall, int nw, const double* __restrict__ wx,
                                                        const double* __restrict__ wy, const double* __restrict__ wz,
                                                        const double* __restrict__ wnx, const double* __restrict__ wny,
                                                        const double* __restrict__ wnz, double* __restrict__ fwd,
                                                        double* __restrict__ fdiss) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    if (is_wall[f]) { fwd[f] = 0.; fdiss[f] = 0.; return; }
    Vec<double> x = mk<double>(fcx[f], fcy[f], fcz[f]);
    double best = 1.e20, nx = 0., ny = 0., nz = 0.;
    for (int q = 0; q < nw; ++q) {
        double d = vmag(vsub(x, mk<double>(wx[q], wy[q], wz[q])));
        if (best > d) { best = d; nx = wnx[q]; ny = wny[q]; nz = wnz[q]; }
    }
    fwd[f] = best;
    fdiss[f] = 1. - fabs(nx * fnx[f] + ny * fny[f] + nz * fnz[f]);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <int FLUX, int VISC, int NV>
static void launch_flux_t(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI,
                          int has_phi, double phi_const, double* FX, int* status, hipStream_t s) {
    hipLaunchKernelGGL((face_flux_kernel<FLUX, VISC, NV>), dim3((g.F + 63) / 64), dim3(64), 0, s, g, P, W, G, PHI, has_phi,
                       phi_const, FX, status);
}

template <int FLUX>
static void launch_flux_v(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI,
                          int has_phi, double phi_const, double* FX, int* status, hipStream_t s) {
    if (g.nv == 6) launch_flux_t<FLUX, 1, 6>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s);   // SA is always viscous
    else if (P.kw) launch_flux_t<FLUX, kViscRans, 5>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s);
    else if (P.viscous) launch_flux_t<FLUX, 1, 5>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s);
    else launch_flux_t<FLUX, 0, 5>(g, P, W, G, PHI, has_phi, phi_const, FX, status, s);
}

template <int FLUX>
static void launch_surface_v(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, int has_phi,
                             double phi_const, double* out, int* status, hipStream_t s) {
    const dim3 grid((g.F + 63) / 64), blk(64);
    if (g.nv == 6) hipLaunchKernelGGL((surface_kernel<FLUX, 1, 6>), grid, blk, 0, s, g, P, W, G, PHI, has_phi, phi_const, out, status);
    else if (P.kw) hipLaunchKernelGGL((surface_kernel<FLUX, kViscRans, 5>), grid, blk, 0, s, g, P, W, G, PHI, has_phi, phi_const, out, status);
    else if (P.viscous) hipLaunchKernelGGL((surface_kernel<FLUX, 1, 5>), grid, blk, 0, s, g, P, W, G, PHI, has_phi, phi_const, out, status);
    else hipLaunchKernelGGL((surface_kernel<FLUX, 0, 5>), grid, blk, 0, s, g, P, W, G, PHI, has_phi, phi_const, out, status);
}

void launch_surface(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, double* out, int* status,
                    hipStream_t s) {
    const int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    const int has_phi = lim != FCFD_LIMITER_NONE;
    const double phi_const = (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0;
    switch (P.flux) {
    case FCFD_FLUX_ROE: launch_surface_v<FCFD_FLUX_ROE>(g, P, W, G, PHI, has_phi, phi_const, out, status, s); break;
    case FCFD_FLUX_VAN_LEER: launch_surface_v<FCFD_FLUX_VAN_LEER>(g, P, W, G, PHI, has_phi, phi_const, out, status, s); break;
    case FCFD_FLUX_SD_SLAU: launch_surface_v<FCFD_FLUX_SD_SLAU>(g, P, W, G, PHI, has_phi, phi_const, out, status, s); break;
    case FCFD_FLUX_SW: launch_surface_v<FCFD_FLUX_SW>(g, P, W, G, PHI, has_phi, phi_const, out, status, s); break;
    default: launch_surface_v<FCFD_FLUX_AUSM_PLUS_UP>(g, P, W, G, PHI, has_phi, phi_const, out, status, s); break;
    }
}

void launch_ghosts(const DevGrid& g, const Params& P, double* W, int* status, hipStream_t s) {
    if (g.B == 0) return;
    if (g.nv == 6) hipLaunchKernelGGL(ghost_kernel<6>, dim3((g.B + 255) / 256), dim3(256), 0, s, g, P, W, status);
    else hipLaunchKernelGGL(ghost_kernel<5>, dim3((g.B + 255) / 256), dim3(256), 0, s, g, P, W, status);
}

template <int NV>
static void launch_vk_range_t(const DevGrid& g, const double* W, double* dref, hipStream_t s) {
    hipLaunchKernelGGL(minmax_partial_kernel<NV>, dim3(kMMBlocks), dim3(256), 0, s, g.N, (size_t)g.ld, g.ref_first, g.ref_last, W,
                       dref + 8);
    hipLaunchKernelGGL(minmax_final_kernel<NV>, dim3(1), dim3(256), 0, s, (size_t)g.ld, g.ref_first, g.ref_last, W, dref + 8, dref);
}

void launch_vk_range(const DevGrid& g, const Params& P, const double* W, double* dref, hipStream_t s) {
    int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    // eps^2 = (1e-6 K dref/L)^2 vanishes for K = 0 (the default threshold,
    // read_inputs.cc:129) whatever dref is, provided L > 0: skip the reduction
    if (lim != FCFD_LIMITER_VK || (P.limiter_threshold == 0.0 && P.glscale > 0.0)) return;
    if (g.nv == 6) launch_vk_range_t<6>(g, W, dref, s);
    else launch_vk_range_t<5>(g, W, dref, s);
}

void launch_green_gauss(const DevGrid& g, const double* W, double* G, int tangent, hipStream_t s, const BcSeed* seed) {
    if (!g.ggc) return;
    BcSeed none{};
    none.index = -1;
    const dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL(gg_kernel<6>, grid, blk, 0, s, g, W, G, tangent, seed ? *seed : none);
    else hipLaunchKernelGGL(gg_kernel<5>, grid, blk, 0, s, g, W, G, tangent, seed ? *seed : none);
}

// face gradients of the wide face-average stencils (Variable<Vec3D>::
// face_calculate, variable.h:134-147): per slot, from 0, entry by entry in
// key order, w times the cell's gradient of u, v, w, T (and nu~); T is the
// value type (the gradients, or a tangent lane's)
template <int NV, typename T>
__global__ void __launch_bounds__(256) wide_grad_kernel(DevGrid g, const T* __restrict__ G, T* __restrict__ out) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= g.nwide) return;
    const size_t ld = g.ld, m = (size_t)g.nwide;
    T acc[3 * (NV - 1)];
#pragma unroll
    for (int k = 0; k < 3 * (NV - 1); ++k) acc[k] = T(0);
    for (int e = g.wptr[sl]; e < g.wptr[sl + 1]; ++e) {
        const int j = g.widx[e];
        const T w = (T)g.ww[e];
#pragma unroll
        for (int k = 0; k < 3 * (NV - 1); ++k) acc[k] = acc[k] + w * G[pidx(ld, 3 + k, j)];
    }
#pragma unroll
    for (int k = 0; k < 3 * (NV - 1); ++k) out[k * m + sl] = acc[k];
}

template <typename T>
static void launch_wide_grad_t(const DevGrid& g, const T* G, T* out, hipStream_t s) {
    if (!g.wslot || g.nwide == 0) return;
    const dim3 grid((g.nwide + 255) / 256), blk(256);
    if (g.nv == 6) hipLaunchKernelGGL((wide_grad_kernel<6, T>), grid, blk, 0, s, g, G, out);
    else hipLaunchKernelGGL((wide_grad_kernel<5, T>), grid, blk, 0, s, g, G, out);
}
void launch_wide_grad(const DevGrid& g, const double* G, double* out, hipStream_t s) { launch_wide_grad_t<double>(g, G, out, s); }
void launch_wide_grad(const DevGrid& g, const float* G, float* out, hipStream_t s) { launch_wide_grad_t<float>(g, G, out, s); }

template <int NV>
static void launch_grad_lim_t(const DevGrid& g, const Params& P, const double* W, double* G, double* PHI, double* dref, double* src,
                              hipStream_t s) {
    int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    launch_green_gauss(g, W, G, 0, s);
    if (lim == FCFD_LIMITER_VK) {
        launch_vk_range(g, P, W, dref, s);
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_VK, NV>), grid, blk, 0, s, g, P, W, G, PHI, dref, src);
    } else if (lim == FCFD_LIMITER_BJ) {
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_BJ, NV>), grid, blk, 0, s, g, P, W, G, PHI, dref, src);
    } else if (lim == FCFD_LIMITER_MINMOD) {
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_MINMOD, NV>), grid, blk, 0, s, g, P, W, G, PHI, dref, src);
    } else {
        hipLaunchKernelGGL((grad_lim_kernel<FCFD_LIMITER_NONE, NV>), grid, blk, 0, s, g, P, W, G, PHI, dref, src);
    }
    launch_wide_grad(g, G, g.wgrad, s);
}

void launch_grad_lim(const DevGrid& g, const Params& P, const double* W, double* G, double* PHI, double* dref, hipStream_t s,
                     double* src) {
    if (g.nv == 6) launch_grad_lim_t<6>(g, P, W, G, PHI, dref, src, s);
    else launch_grad_lim_t<5>(g, P, W, G, PHI, dref, nullptr, s);
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
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL((cell_sum_kernel<6, false, double>), grid, blk, 0, s, g, nullptr, FX, R);
    else hipLaunchKernelGGL((cell_sum_kernel<5, false, double>), grid, blk, 0, s, g, nullptr, FX, R);
}

void launch_cell_sum(const DevGrid& g, const float* FX, float* R, hipStream_t s) {
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL((cell_sum_kernel<6, false, float>), grid, blk, 0, s, g, nullptr, FX, R);
    else hipLaunchKernelGGL((cell_sum_kernel<5, false, float>), grid, blk, 0, s, g, nullptr, FX, R);
}

void launch_residual_sum(const DevGrid& g, const float* src, const float* FX, float* R, hipStream_t s) {
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL((cell_sum_kernel<6, true, float>), grid, blk, 0, s, g, src, FX, R);
    else hipLaunchKernelGGL((cell_sum_kernel<5, false, float>), grid, blk, 0, s, g, nullptr, FX, R);
}

void launch_residual_sum(const DevGrid& g, const double* src, const double* FX, double* R, hipStream_t s) {
    dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL((cell_sum_kernel<6, true, double>), grid, blk, 0, s, g, src, FX, R);
    else hipLaunchKernelGGL((cell_sum_kernel<5, false, double>), grid, blk, 0, s, g, nullptr, FX, R);
}

// CFL_MAX and FIXED time steps and the largest CFL number
// (update_time_step, time_step.cc:155-213).  Minima and maxima are exact, so
// the two-pass reductions give the reference's values whatever their order.
constexpr int kDtBlocks = 256;

__global__ void __launch_bounds__(256) dt_reduce_partial_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                                const double* __restrict__ dt, const double* __restrict__ cur,
                                                                double cur_val, int want_max, double* __restrict__ part) {
    // want_max = 0: min of dt[c]; 1: max of (|V| + a) cur / L with cur = *cur
    // (or cur_val when cur is NULL)
    __shared__ double sh[256];
    const double step = cur ? *cur : cur_val;
    double m = want_max ? -1. : 1.e20;
    const size_t ld = g.ld;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < g.N; c += gridDim.x * 256) {
        if (want_max) {
            double p = W[c], T = W[4 * ld + c];
            Vec<double> V = mk<double>(W[ld + c], W[2 * ld + c], W[3 * ld + c]);
            m = fmax(m, (vmag(V) + m_a(P, p, T)) * step / g.lscale[c]);
        } else {
            m = fmin(m, dt[c]);
        }
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = want_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + w]) : fmin(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) dt_reduce_final_kernel(const double* __restrict__ part, int want_max, double* __restrict__ out) {
    __shared__ double sh[256];
    sh[threadIdx.x] = part[threadIdx.x];   // kDtBlocks == blockDim.x partials
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = want_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + w]) : fmin(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ void __launch_bounds__(256) dt_fill_kernel(int N, const double* __restrict__ src, double val, double* __restrict__ dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < N) dt[c] = src ? *src : val;
}

void launch_time_step(const DevGrid& g, const Params& P, const double* W, int dt_type, double value, double* dt, double* red,
                      hipStream_t s) {
    const dim3 grid((g.N + 255) / 256), blk(256);
    if (dt_type == FCFD_DT_FIXED) {
        hipLaunchKernelGGL(dt_fill_kernel, grid, blk, 0, s, g.N, (const double*)nullptr, value, dt);
        return;
    }
    launch_dt_local(g, P, W, value, dt, s);
    if (dt_type == FCFD_DT_CFL_MAX) {
        hipLaunchKernelGGL(dt_reduce_partial_kernel, dim3(kDtBlocks), blk, 0, s, g, P, W, (const double*)dt, (const double*)nullptr,
                           0.0, 0, red + 2);
        hipLaunchKernelGGL(dt_reduce_final_kernel, dim3(1), blk, 0, s, (const double*)(red + 2), 0, red);
        hipLaunchKernelGGL(dt_fill_kernel, grid, blk, 0, s, g.N, (const double*)red, 0.0, dt);
    }
}

void launch_max_cfl(const DevGrid& g, const Params& P, const double* W, const double* step, double step_val, double* red,
                    hipStream_t s) {
    hipLaunchKernelGGL(dt_reduce_partial_kernel, dim3(kDtBlocks), dim3(256), 0, s, g, P, W, (const double*)nullptr, step, step_val, 1,
                       red + 2);
    hipLaunchKernelGGL(dt_reduce_final_kernel, dim3(1), dim3(256), 0, s, (const double*)(red + 2), 1, red + 1);
}

void launch_store_scalar(const double* src, double val, double* dst, hipStream_t s) {
    hipLaunchKernelGGL(dt_fill_kernel, dim3(1), dim3(64), 0, s, 1, src, val, dst);
}

void launch_dt_local(const DevGrid& g, const Params& P, const double* W, double cfl, double* dt, hipStream_t s) {
    hipLaunchKernelGGL(dt_local_kernel, cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g, P, W, cfl, dt);
}

void launch_wall_distance(int T, const double* ccx, const double* ccy, const double* ccz, int nw, const double* wx, const double* wy,
                          const double* wz, double* wd, hipStream_t s) {
    if (T > 0) hipLaunchKernelGGL(wall_distance_kernel, dim3((T + 255) / 256), dim3(256), 0, s, T, ccx, ccy, ccz, nw, wx, wy, wz, wd);
}

void launch_face_wall(int F, const double* fcx, const double* fcy, const double* fcz, const double* fnx, const double* fny,
                      const double* fnz, const int* is_wall, int nw, const double* wx, const double* wy, const double* wz,
                      const double* wnx, const double* wny, const double* wnz, double* fwd, double* fdiss, hipStream_t s) {
    if (F > 0)
        hipLaunchKernelGGL(face_wall_kernel, dim3((F + 255) / 256), dim3(256), 0, s, F, fcx, fcy, fcz, fnx, fny, fnz, is_wall, nw,
                           wx, wy, wz, wnx, wny, wnz, fwd, fdiss);
}

}  // namespace fcfd
