// Adjoint (vector-Jacobian product) of the second-order residual, and the
// surface loads of ns_assemble_linear_system.cc:109-119.
//
// The reverse sweep is the exact transpose of the FULL tangent
// (fcfd_tangent.hip): ghosts -> gradients -> limiters -> face fluxes -> cell
// sums, walked backwards.  Every stage is a gather in a fixed order (no
// atomics), so the result is deterministic run to run:
//
//   1. face_adj_kernel   per face: the flux adjoint Fbar = Rbar_P - Rbar_N
//                        (+ a load weight on boundary faces) contracted with
//                        the face-flux derivative w.r.t. each of the face's
//                        inputs -- the reconstructed left/right states, the
//                        cell-centre states and the face gradients (20
//                        inputs inviscid, 32 viscous), one forward dual pass
//                        per input through the same physics code -> Z[k][f]
//   2. cell_adj_kernel   per cell: gathers Z over the cell's half-faces and
//                        transposes the MUSCL reconstruction and the face
//                        gradient average -> state, gradient and limiter
//                        adjoints; transposes the cell's limiter (one dual
//                        pass per input direction) -> own-state and gradient
//                        adjoints plus one contribution per neighbour (LH)
//   3. gradt_kernel      per cell and ghost: transposed gradient map gather
//                        and the neighbours' limiter contributions
//   4. ghost_adj_kernel  per cell: transposes the ghost states of its
//                        boundary faces back to the parent
// The VK reference range is held constant, as in the tangent.
#include <hip/hip_runtime.h>

#include "fcfd_adjoint.h"
#include "fcfd_face.h"
#include "fcfd_limiter.h"

namespace fcfd {

namespace {

__device__ __forceinline__ Dual seed(double v, int k, int idx) { return Dual(v, k == idx ? 1.0 : 0.0); }

// Z rows, NV variables per state: [0, NV) left face state, [NV, 2NV) left
// centre, [2NV, 3NV) right face state (interior faces), [3NV, 4NV) right
// centre (ghost centre on a boundary face), then the face gradients of
// u, v, w, T (and nu~) when viscous
template <int NV> struct ZRows {
    static constexpr int L = 0, Lc = NV, R = 2 * NV, Rc = 3 * NV, G = 4 * NV;
    static constexpr int ngrad = NV == 6 ? 15 : 12;
};
template <int VISC, int NV> constexpr int zcount() { return 4 * NV + (VISC ? ZRows<NV>::ngrad : 0); }

template <int FLUX, int VISC, int NV>
__global__ void __launch_bounds__(256) face_adj_kernel(DevGrid g, Params P, CellLoader<NV> cl, const double* __restrict__ Rbar,
                                                       const double* __restrict__ fxw, double* __restrict__ Z,
                                                       int* __restrict__ status) {
    using Zr = ZRows<NV>;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    const size_t Fs = g.F, ld = g.ld;
    const int4 cr = g.fcell[f];
    const int p = cr.x, n = cr.y, bc = cr.z;
    const bool bnd = bc >= 0;
    constexpr int nz = zcount<VISC, NV>();
    double fb[NV];
    bool any = false;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        fb[v] = Rbar[v * ld + p] - (bnd ? 0.0 : Rbar[v * ld + n]);
        if (bnd && fxw && v < 5) fb[v] += fxw[5 * bc + v];
        any |= fb[v] != 0.0;
    }
    if (!any) {
        for (int k = 0; k < nz; ++k) Z[k * Fs + f] = 0.0;
        return;
    }
    FaceCtx<double, NV> X0;
    face_inputs<VISC, NV>(g, cl, P, f, X0);
#pragma unroll 1
    for (int k = 0; k < nz; ++k) {
        if (bnd && k >= Zr::R && k < Zr::Rc) { Z[k * Fs + f] = 0.0; continue; }
        FaceCtx<Dual, NV> X;
        X.F = X0.F; X.p = p; X.n = n;
        if constexpr (VISC == kViscRans) X.mut_face = X0.mut_face;
        X.L.p = seed(X0.L.p, k, Zr::L);
        X.L.V = mk<Dual>(seed(X0.L.V.x, k, Zr::L + 1), seed(X0.L.V.y, k, Zr::L + 2), seed(X0.L.V.z, k, Zr::L + 3));
        X.L.T = seed(X0.L.T, k, Zr::L + 4);
        X.L.p_center = seed(X0.L.p_center, k, Zr::Lc);
        X.L.V_center = mk<Dual>(seed(X0.L.V_center.x, k, Zr::Lc + 1), seed(X0.L.V_center.y, k, Zr::Lc + 2),
                                seed(X0.L.V_center.z, k, Zr::Lc + 3));
        X.L.T_center = seed(X0.L.T_center, k, Zr::Lc + 4);
        if constexpr (NV == 6) {
            X.L.nut = seed(X0.L.nut, k, Zr::L + 5);
            X.L.nut_center = seed(X0.L.nut_center, k, Zr::Lc + 5);
        }
        X.L.rho = m_rho(P, X.L.p, X.L.T);
        {
            CState<Dual> Rs;
            if (!bnd) {
                Rs.p = seed(X0.R.p, k, Zr::R);
                Rs.V = mk<Dual>(seed(X0.R.V.x, k, Zr::R + 1), seed(X0.R.V.y, k, Zr::R + 2), seed(X0.R.V.z, k, Zr::R + 3));
                Rs.T = seed(X0.R.T, k, Zr::R + 4);
                if constexpr (NV == 6) Rs.nut = seed(X0.R.nut, k, Zr::R + 5);
                Rs.rho = m_rho(P, Rs.p, Rs.T);
            }
            Rs.p_center = seed(X0.R.p_center, k, Zr::Rc);
            Rs.V_center = mk<Dual>(seed(X0.R.V_center.x, k, Zr::Rc + 1), seed(X0.R.V_center.y, k, Zr::Rc + 2),
                                   seed(X0.R.V_center.z, k, Zr::Rc + 3));
            Rs.T_center = seed(X0.R.T_center, k, Zr::Rc + 4);
            if constexpr (NV == 6) Rs.nut_center = seed(X0.R.nut_center, k, Zr::Rc + 5);
            X.R = Rs;
        }
        if (VISC) {
            X.fs.gu = mk<Dual>(seed(X0.fs.gu.x, k, Zr::G + 0), seed(X0.fs.gu.y, k, Zr::G + 1), seed(X0.fs.gu.z, k, Zr::G + 2));
            X.fs.gv = mk<Dual>(seed(X0.fs.gv.x, k, Zr::G + 3), seed(X0.fs.gv.y, k, Zr::G + 4), seed(X0.fs.gv.z, k, Zr::G + 5));
            X.fs.gw = mk<Dual>(seed(X0.fs.gw.x, k, Zr::G + 6), seed(X0.fs.gw.y, k, Zr::G + 7), seed(X0.fs.gw.z, k, Zr::G + 8));
            X.fs.gT = mk<Dual>(seed(X0.fs.gT.x, k, Zr::G + 9), seed(X0.fs.gT.y, k, Zr::G + 10), seed(X0.fs.gT.z, k, Zr::G + 11));
            if constexpr (NV == 6)
                X.fs.gn = mk<Dual>(seed(X0.fs.gn.x, k, Zr::G + 12), seed(X0.fs.gn.y, k, Zr::G + 13), seed(X0.fs.gn.z, k, Zr::G + 14));
        }
        face_tail<FLUX, VISC, NV>(g, P, f, X, status);
        double zb = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) zb += fb[v] * (X.D[v].d - X.C[v].d);
        Z[k * Fs + f] = zb;
    }
}

// neighbour states for the limiter, seeding the one across half-face `dir`
struct AdjNb {
    const double* __restrict__ W;
    size_t ld;
    int dir;
    __device__ __forceinline__ Dual nb(int h, int j, int v) const { return Dual(W[v * ld + j], h == dir ? 1.0 : 0.0); }
};

template <int LIM, int VISC, int NV>
__global__ void __launch_bounds__(256) cell_adj_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                       const double* __restrict__ G, const double* __restrict__ PHI,
                                                       double phi_const, const double* __restrict__ dref,
                                                       const double* __restrict__ Rbar, const double* __restrict__ Z,
                                                       double* __restrict__ Wb1, double* __restrict__ Gb, double* __restrict__ LH) {
    using Zr = ZRows<NV>;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N + g.B) return;
    const size_t Fs = g.F, ld = g.ld, nhf = g.nhf;
    if (c >= g.N) {   // a ghost's centre state enters its face only
        const int f = g.gh_face[c - g.N];
#pragma unroll
        for (int v = 0; v < NV; ++v) Wb1[v * ld + c] = Z[(Zr::Rc + v) * Fs + f];
        return;
    }
    constexpr bool has_phi = LIM != FCFD_LIMITER_NONE;
    double qb[NV], gb[3 * NV], pb[NV], gc[3 * NV], ph[NV];
    load_pairs<3 * NV>(G, ld, c, gc);
#pragma unroll
    for (int k = 0; k < 3 * NV; ++k) gb[k] = 0.0;
    if (has_phi) load_pairs<NV>(PHI, ld, c, ph);
#pragma unroll
    for (int v = 0; v < NV; ++v) { qb[v] = 0.0; pb[v] = 0.0; if (!has_phi) ph[v] = phi_const; }
    const Vec<double> xc = centroid(g, c);
    const int h0 = g.hf_ptr[c], h1 = g.hf_ptr[c + 1];
    for (int h = h0; h < h1; ++h) {
        const int fs = g.hf_face[h];
        const bool left = fs >= 0;
        const int f = left ? fs : ~fs;
        const int of = left ? Zr::L : Zr::R, oc = left ? Zr::Lc : Zr::Rc;
        const Vec<double> c2f = vsub(face_centroid(g, f), xc);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            // q_f = q_c + phi (grad q . c2f)  (recon) and the centre value itself
            const double zf = Z[(of + v) * Fs + f], zc = Z[(oc + v) * Fs + f];
            qb[v] += zf;
            qb[v] += zc;
            pb[v] += zf * (gc[3 * v] * c2f.x + gc[3 * v + 1] * c2f.y + gc[3 * v + 2] * c2f.z);
            const double s = zf * ph[v];
            gb[3 * v] += s * c2f.x;
            gb[3 * v + 1] += s * c2f.y;
            gb[3 * v + 2] += s * c2f.z;
        }
        if (VISC && !(g.wslot && g.wslot[f] >= 0)) {   // face gradient = w_P grad_P + w_N grad_N (interior), grad_P (boundary)
            const double2 wpn = face_pair(g, 5, f);
            const double w = g.fbc[f] >= 0 ? 1.0 : (left ? wpn.x : wpn.y);
#pragma unroll
            for (int m = 0; m < 12; ++m) gb[3 + m] += w * Z[(Zr::G + m) * Fs + f];
            if constexpr (NV == 6) {
#pragma unroll
                for (int m = 0; m < 3; ++m) gb[15 + m] += w * Z[(Zr::G + 12 + m) * Fs + f];
            }
        }
    }
    if (VISC && g.wtptr) {   // wide stencils: this cell's weight in each of their face gradients
        for (int e = g.wtptr[c]; e < g.wtptr[c + 1]; ++e) {
            const int f = g.wtface[e];
            const double w = g.wtw[e];
#pragma unroll
            for (int m = 0; m < Zr::ngrad; ++m) gb[3 + m] += w * Z[(Zr::G + m) * Fs + f];
        }
    }
    if constexpr (NV == 6) {   // SA source V S(p, T, nu~, grad u, grad v, grad w, grad nu~)
        const double rb = Rbar[5 * ld + c] * g.vol[c];
        if (rb != 0.0) {
            const double d = g.wd[c];
            // directions 0..2: p, T, nu~; 3..14: gradient components of u, v, w, nu~
            // (unrolled: the accumulation targets are then compile-time registers)
#pragma unroll
            for (int k = 0; k < 15; ++k) {
                Dual p(W[c], k == 0 ? 1.0 : 0.0), T(W[4 * ld + c], k == 1 ? 1.0 : 0.0), nut(W[5 * ld + c], k == 2 ? 1.0 : 0.0);
                auto gvec = [&](int v, int base) {
                    Vec<Dual> r;
                    r.x = Dual(gc[3 * v], k == base ? 1.0 : 0.0);
                    r.y = Dual(gc[3 * v + 1], k == base + 1 ? 1.0 : 0.0);
                    r.z = Dual(gc[3 * v + 2], k == base + 2 ? 1.0 : 0.0);
                    return r;
                };
                Dual S = sa_source<Dual>(m_rho(P, p, T), m_visc(P, T), nut, gvec(1, 3), gvec(2, 6), gvec(3, 9), gvec(5, 12), d);
                const double a = rb * S.d;
                if (k == 0) qb[0] += a;
                else if (k == 1) qb[4] += a;
                else if (k == 2) qb[5] += a;
                else if (k < 12) gb[3 + (k - 3)] += a;
                else gb[15 + (k - 12)] += a;
            }
        }
    }
    if (has_phi) {
        bool any = false;
#pragma unroll
        for (int v = 0; v < NV; ++v) any |= pb[v] != 0.0;
        if (!any) {
            for (int h = h0; h < h1; ++h)
#pragma unroll
                for (int v = 0; v < NV; ++v) LH[v * nhf + h] = 0.0;
        } else {
            // directions: -1 own state, -2..-4 own gradient x/y/z, h >= h0 the
            // neighbour across half-face h
#pragma unroll 1
            for (int dir = -4; dir < h1; dir = (dir == -1) ? h0 : dir + 1) {
                Dual q0[NV], gr[3 * NV], phi[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    q0[v] = Dual(W[v * ld + c], dir == -1 ? 1.0 : 0.0);
                    gr[3 * v] = Dual(gc[3 * v], dir == -2 ? 1.0 : 0.0);
                    gr[3 * v + 1] = Dual(gc[3 * v + 1], dir == -3 ? 1.0 : 0.0);
                    gr[3 * v + 2] = Dual(gc[3 * v + 2], dir == -4 ? 1.0 : 0.0);
                }
                limiter_phi<LIM, NV>(g, P, AdjNb{W, ld, dir}, c, q0, gr, dref, phi);
                if (dir == -1) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) qb[v] += pb[v] * phi[v].d;
                } else if (dir < -1) {
                    const int k = -2 - dir;
#pragma unroll
                    for (int v = 0; v < NV; ++v) gb[3 * v + k] += pb[v] * phi[v].d;
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) LH[v * nhf + dir] = pb[v] * phi[v].d;
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) Wb1[v * ld + c] = qb[v];
#pragma unroll
    for (int k = 0; k < 3 * NV; ++k) Gb[k * ld + c] = gb[k];
}

// transposed gradient map + the neighbours' limiter contributions
template <int NV>
__global__ void __launch_bounds__(256) gradt_kernel(DevGrid g, AdjGrid a, int has_lim, const double* __restrict__ Wb1,
                                                    const double* __restrict__ Gb, const double* __restrict__ LH,
                                                    double* __restrict__ Wb2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.N + g.B) return;
    const size_t ld = g.ld, nhf = g.nhf;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = Wb1[v * ld + j];
    for (int e = a.tg_ptr[j]; e < a.tg_ptr[j + 1]; ++e) {
        const int c = a.tg_cell[e];
        const double wx = a.tgx[e], wy = a.tgy[e], wz = a.tgz[e];
        const int fixed = a.tgm ? a.tgm[e] : 0;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            if (!((fixed >> v) & 1)) acc[v] += wx * Gb[(3 * v) * ld + c] + wy * Gb[(3 * v + 1) * ld + c] + wz * Gb[(3 * v + 2) * ld + c];
    }
    if (has_lim) {
        if (j < g.N) {
            for (int h = g.hf_ptr[j]; h < g.hf_ptr[j + 1]; ++h) {
                const int t = a.hf_twin[h];
                if (t < 0) continue;
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] += LH[v * nhf + t];
            }
        } else {
            const int t = a.gh_hf[j - g.N];
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += LH[v * nhf + t];
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) Wb2[v * ld + j] = acc[v];
}

// ghost state of boundary face f from its parent c (update_boundaries as in
// ghost_kernel), with a unit tangent on parent variable m
template <int NV>
__device__ __forceinline__ void ghost_dual(const DevGrid& g, const Params& P, int f, int c, const double* __restrict__ W, int m,
                                           Dual out[NV], int* status) {
    FaceGeo F;
    load_geo(g, f, F);
    const size_t ld = g.ld;
    CState<Dual> L;
    L.p = L.p_center = Dual(W[c], m == 0 ? 1.0 : 0.0);
    L.V = L.V_center = mk<Dual>(Dual(W[ld + c], m == 1 ? 1.0 : 0.0), Dual(W[2 * ld + c], m == 2 ? 1.0 : 0.0),
                                Dual(W[3 * ld + c], m == 3 ? 1.0 : 0.0));
    L.T = L.T_center = Dual(W[4 * ld + c], m == 4 ? 1.0 : 0.0);
    if constexpr (NV == 6) L.nut = L.nut_center = Dual(W[5 * ld + c], m == 5 ? 1.0 : 0.0);
    L.rho = m_rho(P, L.p, L.T);
    finish_state(P, L, F);
    CState<Dual> R = L;
    if (apply_bcs(P, g.bc[F.bc], L, R, F)) atomicOr(status, 1);
    out[0] = R.p_center;
    out[1] = R.V_center.x;
    out[2] = R.V_center.y;
    out[3] = R.V_center.z;
    out[4] = R.T_center;
    if constexpr (NV == 6) out[5] = R.nut_center;
}

template <int NV>
__global__ void __launch_bounds__(256) ghost_adj_kernel(DevGrid g, Params P, const double* __restrict__ W,
                                                        const double* __restrict__ Wb2, double* __restrict__ Wbar,
                                                        int* __restrict__ status) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = Wb2[v * ld + c];
    for (int h = g.hf_ptr[c]; h < g.hf_ptr[c + 1]; ++h) {
        const int f = g.hf_face[h];
        if (f < 0 || g.fbc[f] < 0) continue;
        const int gi = g.fnei[f];
        double wg[NV];
        bool any = false;
#pragma unroll
        for (int v = 0; v < NV; ++v) { wg[v] = Wb2[v * ld + gi]; any |= wg[v] != 0.0; }
        if (!any) continue;
#pragma unroll 1
        for (int m = 0; m < NV; ++m) {
            Dual o[NV];
            ghost_dual<NV>(g, P, f, c, W, m, o, status);
            double s = 0.0;
#pragma unroll
            for (int v = 0; v < NV; ++v) s += wg[v] * o[v].d;
            acc[m] += s;
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) Wbar[v * ld + c] = acc[v];
}

// ---- surface loads -----------------------------------------------------------
template <int FLUX, int VISC, int NV>
__global__ void __launch_bounds__(256) loads_face_kernel(DevGrid g, AdjGrid a, Params P, CellLoader<NV> cl, double mcx, double mcy,
                                                         double mcz, double* __restrict__ tmp, int nlf, int* __restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlf) return;
    const int f = a.lf_face[i];
    FaceCtx<double, NV> X;
    face_base<FLUX, VISC, NV>(g, P, cl, f, X, status);
    // force += (C - D)[1..3], moment += (x_f - x_m) x force (:109-119)
    Vec<double> t = mk<double>(X.C[1] - X.D[1], X.C[2] - X.D[2], X.C[3] - X.D[3]);
    const Vec<double> xf = face_centroid(g, f);
    Vec<double> r = mk<double>(xf.x - mcx, xf.y - mcy, xf.z - mcz);
    Vec<double> mm = vcross(r, t);
    const size_t s = nlf;
    tmp[i] = t.x; tmp[s + i] = t.y; tmp[2 * s + i] = t.z;
    tmp[3 * s + i] = mm.x; tmp[4 * s + i] = mm.y; tmp[5 * s + i] = mm.z;
}

// one thread per (BC, component): the reference's sequential face-order sum
__global__ void __launch_bounds__(64) loads_sum_kernel(AdjGrid a, const double* __restrict__ tmp, int nlf, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * a.n_bc) return;
    const int b = t / 6, k = t % 6;
    double acc = 0.0;
    for (int i = a.lf_ptr[b]; i < a.lf_ptr[b + 1]; ++i) acc += tmp[(size_t)k * nlf + i];
    out[t] = acc;
}

// ---- dispatch -----------------------------------------------------------------
template <int NV>
struct AdjLaunch {
    template <int FLUX, int VISC>
    static void face(const DevGrid& g, const Params& P, const CellLoader<NV>& cl, const double* Rbar, const double* fxw, double* Z,
                     int* status, hipStream_t s) {
        hipLaunchKernelGGL((face_adj_kernel<FLUX, VISC, NV>), dim3((g.F + 255) / 256), dim3(256), 0, s, g, P, cl, Rbar, fxw, Z, status);
    }
    template <int FLUX>
    static void face_v(const DevGrid& g, const Params& P, const CellLoader<NV>& cl, const double* Rbar, const double* fxw, double* Z,
                       int* status, hipStream_t s) {
        if constexpr (NV == 6) face<FLUX, 1>(g, P, cl, Rbar, fxw, Z, status, s);
        else if (P.kw) face<FLUX, kViscRans>(g, P, cl, Rbar, fxw, Z, status, s);
        else if (P.viscous) face<FLUX, 1>(g, P, cl, Rbar, fxw, Z, status, s);
        else face<FLUX, 0>(g, P, cl, Rbar, fxw, Z, status, s);
    }
    template <int LIM>
    static void cell(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, double phi_const,
                     const double* dref, const double* Rbar, const AdjWork& w, hipStream_t s) {
        dim3 grid((g.N + g.B + 255) / 256), blk(256);
        if constexpr (NV == 6)
            hipLaunchKernelGGL((cell_adj_kernel<LIM, 1, NV>), grid, blk, 0, s, g, P, W, G, PHI, phi_const, dref, Rbar, w.Z, w.Wb1, w.Gb, w.LH);
        else if (P.viscous)
            hipLaunchKernelGGL((cell_adj_kernel<LIM, 1, NV>), grid, blk, 0, s, g, P, W, G, PHI, phi_const, dref, Rbar, w.Z, w.Wb1, w.Gb, w.LH);
        else
            hipLaunchKernelGGL((cell_adj_kernel<LIM, 0, NV>), grid, blk, 0, s, g, P, W, G, PHI, phi_const, dref, Rbar, w.Z, w.Wb1, w.Gb, w.LH);
    }
    template <int FLUX>
    static void loads_v(const DevGrid& g, const AdjGrid& a, const Params& P, const CellLoader<NV>& cl, double mcx, double mcy, double mcz,
                        double* tmp, int nlf, int* status, hipStream_t s) {
        dim3 grid((nlf + 255) / 256), blk(256);
        if constexpr (NV == 6)
            hipLaunchKernelGGL((loads_face_kernel<FLUX, 1, NV>), grid, blk, 0, s, g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status);
        else if (P.kw)
            hipLaunchKernelGGL((loads_face_kernel<FLUX, kViscRans, NV>), grid, blk, 0, s, g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status);
        else if (P.viscous)
            hipLaunchKernelGGL((loads_face_kernel<FLUX, 1, NV>), grid, blk, 0, s, g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status);
        else
            hipLaunchKernelGGL((loads_face_kernel<FLUX, 0, NV>), grid, blk, 0, s, g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status);
    }

    static void vjp(const DevGrid& g, const AdjGrid& a, const Params& P, double* W, double* G, double* PHI, double* dref,
                    const double* Rbar, const double* fxw, const AdjWork& w, double* Wbar, int* status, hipStream_t s) {
        const int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
        const bool has_phi = lim != FCFD_LIMITER_NONE;
        const double phi_const = (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0;
        launch_ghosts(g, P, W, status, s);
        launch_grad_lim(g, P, W, G, PHI, dref, s);
        CellLoader<NV> cl{W, G, PHI, has_phi, phi_const, (size_t)g.ld};
        switch (P.flux) {
        case FCFD_FLUX_ROE: face_v<FCFD_FLUX_ROE>(g, P, cl, Rbar, fxw, w.Z, status, s); break;
        case FCFD_FLUX_VAN_LEER: face_v<FCFD_FLUX_VAN_LEER>(g, P, cl, Rbar, fxw, w.Z, status, s); break;
        case FCFD_FLUX_SD_SLAU: face_v<FCFD_FLUX_SD_SLAU>(g, P, cl, Rbar, fxw, w.Z, status, s); break;
        case FCFD_FLUX_SW: face_v<FCFD_FLUX_SW>(g, P, cl, Rbar, fxw, w.Z, status, s); break;
        default: face_v<FCFD_FLUX_AUSM_PLUS_UP>(g, P, cl, Rbar, fxw, w.Z, status, s); break;
        }
        switch (lim) {
        case FCFD_LIMITER_VK: cell<FCFD_LIMITER_VK>(g, P, W, G, PHI, phi_const, dref, Rbar, w, s); break;
        case FCFD_LIMITER_BJ: cell<FCFD_LIMITER_BJ>(g, P, W, G, PHI, phi_const, dref, Rbar, w, s); break;
        case FCFD_LIMITER_MINMOD: cell<FCFD_LIMITER_MINMOD>(g, P, W, G, PHI, phi_const, dref, Rbar, w, s); break;
        default: cell<FCFD_LIMITER_NONE>(g, P, W, G, PHI, phi_const, dref, Rbar, w, s); break;
        }
        hipLaunchKernelGGL(gradt_kernel<NV>, dim3((g.N + g.B + 255) / 256), dim3(256), 0, s, g, a, has_phi ? 1 : 0, w.Wb1, w.Gb, w.LH, w.Wb2);
        hipLaunchKernelGGL(ghost_adj_kernel<NV>, dim3((g.N + 255) / 256), dim3(256), 0, s, g, P, W, w.Wb2, Wbar, status);
    }

    static void loads(const DevGrid& g, const AdjGrid& a, const Params& P, const double* W, const double* G, const double* PHI,
                      double mcx, double mcy, double mcz, double* tmp, int nlf, int* status, hipStream_t s) {
        const int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
        CellLoader<NV> cl{W, G, PHI, lim != FCFD_LIMITER_NONE, (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0, (size_t)g.ld};
        if (nlf <= 0) return;
        switch (P.flux) {
        case FCFD_FLUX_ROE: loads_v<FCFD_FLUX_ROE>(g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status, s); break;
        case FCFD_FLUX_VAN_LEER: loads_v<FCFD_FLUX_VAN_LEER>(g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status, s); break;
        case FCFD_FLUX_SD_SLAU: loads_v<FCFD_FLUX_SD_SLAU>(g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status, s); break;
        case FCFD_FLUX_SW: loads_v<FCFD_FLUX_SW>(g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status, s); break;
        default: loads_v<FCFD_FLUX_AUSM_PLUS_UP>(g, a, P, cl, mcx, mcy, mcz, tmp, nlf, status, s); break;
        }
    }
};

}  // namespace

void launch_vjp(const DevGrid& g, const AdjGrid& a, const Params& P, double* W, double* G, double* PHI, double* dref,
                const double* Rbar, const double* fxw, const AdjWork& w, double* Wbar, int* status, hipStream_t s) {
    if (g.nv == 6) AdjLaunch<6>::vjp(g, a, P, W, G, PHI, dref, Rbar, fxw, w, Wbar, status, s);
    else AdjLaunch<5>::vjp(g, a, P, W, G, PHI, dref, Rbar, fxw, w, Wbar, status, s);
}

// the loads' tangent from the faces' d(D - C) (dFX [nv][F]): per load face
// dt = -dFX[1..3], dm = (x_f - x_m) x dt, then the same ordered per-BC sums
__global__ void __launch_bounds__(256) loads_tangent_face_kernel(DevGrid g, AdjGrid a, const double* __restrict__ dFX, double mcx,
                                                                 double mcy, double mcz, double* __restrict__ tmp, int nlf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlf) return;
    const int f = a.lf_face[i];
    const size_t Fs = g.F;
    Vec<double> t = mk<double>(-dFX[Fs + f], -dFX[2 * Fs + f], -dFX[3 * Fs + f]);
    const Vec<double> xf = face_centroid(g, f);
    Vec<double> r = mk<double>(xf.x - mcx, xf.y - mcy, xf.z - mcz);
    Vec<double> mm = vcross(r, t);
    const size_t s = nlf;
    tmp[i] = t.x; tmp[s + i] = t.y; tmp[2 * s + i] = t.z;
    tmp[3 * s + i] = mm.x; tmp[4 * s + i] = mm.y; tmp[5 * s + i] = mm.z;
}

void launch_loads_tangent(const DevGrid& g, const AdjGrid& a, const double* dFX, double mcx, double mcy, double mcz, double* tmp,
                          int nlf, double* out, hipStream_t s) {
    if (nlf > 0) hipLaunchKernelGGL(loads_tangent_face_kernel, dim3((nlf + 255) / 256), dim3(256), 0, s, g, a, dFX, mcx, mcy, mcz, tmp, nlf);
    if (a.n_bc > 0) hipLaunchKernelGGL(loads_sum_kernel, dim3((6 * a.n_bc + 63) / 64), dim3(64), 0, s, a, tmp, nlf, out);
}

// sum over boundary faces of sum_v w[5 bc + v] dFX[v][f] (v < 5): the
// boundary-flux part of an output's tangent; partials per block, then a
// fixed-order tree (part: 256 doubles, out: 1)
constexpr int kBfBlocks = 256;
__global__ void __launch_bounds__(256) bface_dot_kernel(DevGrid g, const double* __restrict__ w, const double* __restrict__ dFX,
                                                        double* __restrict__ part) {
    __shared__ double sh[256];
    double acc = 0.0;
    const size_t Fs = g.F;
    for (int f = blockIdx.x * 256 + threadIdx.x; f < g.F; f += kBfBlocks * 256) {
        const int b = g.fcell[f].z;
        if (b < 0) continue;
#pragma unroll
        for (int v = 0; v < 5; ++v) acc += w[5 * b + v] * dFX[v * Fs + f];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int q = 128; q > 0; q >>= 1) {
        if (threadIdx.x < q) sh[threadIdx.x] += sh[threadIdx.x + q];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(256) bface_dot_final_kernel(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[256];
    sh[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int q = 128; q > 0; q >>= 1) {
        if (threadIdx.x < q) sh[threadIdx.x] += sh[threadIdx.x + q];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

void launch_bface_dot(const DevGrid& g, const double* w, const double* dFX, double* part, double* out, hipStream_t s) {
    hipLaunchKernelGGL(bface_dot_kernel, dim3(kBfBlocks), dim3(256), 0, s, g, w, dFX, part);
    hipLaunchKernelGGL(bface_dot_final_kernel, dim3(1), dim3(256), 0, s, part, out);
}

void launch_loads(const DevGrid& g, const AdjGrid& a, const Params& P, const double* W, const double* G, const double* PHI,
                  double mcx, double mcy, double mcz, double* tmp, int nlf, double* out, int* status, hipStream_t s) {
    if (g.nv == 6) AdjLaunch<6>::loads(g, a, P, W, G, PHI, mcx, mcy, mcz, tmp, nlf, status, s);
    else AdjLaunch<5>::loads(g, a, P, W, G, PHI, mcx, mcy, mcz, tmp, nlf, status, s);
    if (a.n_bc > 0) hipLaunchKernelGGL(loads_sum_kernel, dim3((6 * a.n_bc + 63) / 64), dim3(64), 0, s, a, tmp, nlf, out);
}

}  // namespace fcfd
