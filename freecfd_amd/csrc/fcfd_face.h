// Per-face state assembly shared by the residual, Jacobian and tangent
// kernels: loads the face geometry and both cell states, reconstructs the
// face states (MUSCL), applies the boundary condition, and evaluates the
// convective and diffusive fluxes, exactly as one iteration of the face loop
// of ns_assemble_linear_system.cc:75-93.  Templated on the scalar type S and
// on a cell loader, so that the tangent kernels run the same code on dual
// numbers.
#pragma once
#include <type_traits>
#include "fcfd_device.h"
#include "fcfd_dual.h"
#include "fcfd_physics.h"

namespace fcfd {

__device__ __forceinline__ double2 face_pair(const DevGrid& g, int row, int f) { return g.fgeo[(size_t)row * g.F + f]; }

__device__ __forceinline__ Vec<double> face_centroid(const DevGrid& g, int f) {
    const double2 a = face_pair(g, 3, f), b = face_pair(g, 4, f);
    return mk<double>(a.y, b.x, b.y);
}

__device__ __forceinline__ void load_geo(const DevGrid& g, int f, FaceGeo& F) {
    const double2 a0 = face_pair(g, 0, f), a1 = face_pair(g, 1, f), a2 = face_pair(g, 2, f), a3 = face_pair(g, 3, f);
    F.n = mk<double>(a0.x, a0.y, a1.x);
    F.t1 = mk<double>(a1.y, a2.x, a2.y);
    F.t2 = vnorm(vcross(F.n, F.t1));
    F.area = a3.x;
    F.bc = g.fcell[f].z;
}

__device__ __forceinline__ Vec<double> centroid(const DevGrid& g, int c) {
    const double2 a = g.cc2[c], b = g.cc2[(size_t)g.ld + c];
    return mk<double>(a.x, a.y, b.x);
}

// primal cell loader: W [NV][ld], gradients and limiters in row pairs
template <int NV>
struct CellLoader {
    const double* __restrict__ W;
    const double* __restrict__ G;
    const double* __restrict__ PHI;
    bool has_phi;
    double phi_const;
    size_t ld;
    __device__ __forceinline__ void load(int c, bool ghost, double q[NV], Vec<double> gr[NV], double ph[NV]) const {
        double g[3 * NV], pl[NV];
        load_pairs<3 * NV>(G, ld, c, g);
        if (has_phi && !ghost) load_pairs<NV>(PHI, ld, c, pl);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            q[v] = W[v * ld + c];
            gr[v] = mk<double>(g[3 * v], g[3 * v + 1], g[3 * v + 2]);
            ph[v] = (has_phi && !ghost) ? pl[v] : (ghost ? 0.0 : phi_const);
        }
    }
    // face gradient of variable 1 + v (u, v, w, T, nu~) of wide-stencil slot sl
    __device__ __forceinline__ Vec<double> face_grad(const DevGrid& g, int sl, int v) const {
        const size_t m = (size_t)g.nwide;
        return mk<double>(g.wgrad[(3 * v) * m + sl], g.wgrad[(3 * v + 1) * m + sl], g.wgrad[(3 * v + 2) * m + sl]);
    }
};

// tangent cell loader: value = primal arrays, derivative = tangent arrays
// (of the derivative lane's type TD).  FULL = false: gradients, limiters and
// ghost states carry no tangent (the reference FD Jacobian's frozen-gradient
// linearisation).
template <int NV, bool FULL, typename TD = double>
struct DualLoader {
    using D = DualT<TD>;
    CellLoader<NV> p;
    const TD* __restrict__ dW;
    const TD* __restrict__ dG;
    const TD* __restrict__ dPHI;
    int N;
    __device__ __forceinline__ void load(int c, bool ghost, D q[NV], Vec<D> gr[NV], D ph[NV]) const {
        double qv[NV], phv[NV];
        Vec<double> gv[NV];
        p.load(c, ghost, qv, gv, phv);
        const size_t ld = p.ld;
        TD dg[3 * NV], dp[NV];
        if (FULL) {
            load_pairs<3 * NV>(dG, ld, c, dg);
            if (p.has_phi && !ghost) load_pairs<NV>(dPHI, ld, c, dp);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            TD dq = (FULL || c < N) ? dW[v * ld + c] : TD(0);
            q[v] = D(qv[v], dq);
            if (FULL) {
                gr[v] = mk<D>(D(gv[v].x, dg[3 * v]), D(gv[v].y, dg[3 * v + 1]), D(gv[v].z, dg[3 * v + 2]));
                ph[v] = D(phv[v], (p.has_phi && !ghost) ? dp[v] : TD(0));
            } else {
                gr[v] = mk<D>(D(gv[v].x), D(gv[v].y), D(gv[v].z));
                ph[v] = D(phv[v]);
            }
        }
    }
    // wide-stencil face gradient of variable 1 + v: the value from the
    // gradient pass, the tangent (FULL) from the tangent's pass over dG
    __device__ __forceinline__ Vec<D> face_grad(const DevGrid& g, int sl, int v) const {
        const Vec<double> a = p.face_grad(g, sl, v);
        if (!FULL) return mk<D>(D(a.x), D(a.y), D(a.z));
        const size_t m = (size_t)g.nwide;
        const TD* dg;
        if constexpr (std::is_same<TD, float>::value) dg = g.wgrad_f;
        else dg = g.wgrad_d;
        return mk<D>(D(a.x, dg[(3 * v) * m + sl]), D(a.y, dg[(3 * v + 1) * m + sl]), D(a.z, dg[(3 * v + 2) * m + sl]));
    }
};

// A face's packed records, all issued before any is used: the cell loads
// depend on the parent/neighbour ids, and issuing the geometry with them
// (rather than after) keeps one round of load latency per face
template <int VISC>
struct FaceRec {
    int4 cr;
    double2 a[6];
    __device__ __forceinline__ void load(const DevGrid& g, int f) {
        cr = g.fcell[f];
#pragma unroll
        for (int r = 0; r < 5; ++r) a[r] = face_pair(g, r, f);
        a[5] = VISC ? face_pair(g, 5, f) : make_double2(0., 0.);
    }
};

// everything one face-loop iteration computes before the residual insertion
template <typename S, int NV>
struct FaceCtx {
    FaceGeo F;
    CState<S> L, R;
    FState<S> fs;
    S C[NV], D[NV];
    int p, n;
    double mut_face;   // VISC = kViscRans: mu_t.face of the two-equation model (frozen: no tangent)
};

// Stage 1 of a face-loop iteration: geometry, the two cell loads and the
// reconstruction.  Leaves X.L reconstructed (rho set, not finished); X.R
// reconstructed for an interior face, or holding only the ghost's centre
// values for a boundary face; X.fs.g* the face gradients when VISC.
template <int VISC, int NV, typename S, typename Loader>
__device__ __forceinline__ void face_inputs(const DevGrid& g, const Loader& ld_, const Params& P, int f, FaceCtx<S, NV>& X) {
    FaceRec<VISC> rec;
    rec.load(g, f);
    FaceGeo& F = X.F;
    F.n = mk<double>(rec.a[0].x, rec.a[0].y, rec.a[1].x);
    F.t1 = mk<double>(rec.a[1].y, rec.a[2].x, rec.a[2].y);
    F.t2 = vnorm(vcross(F.n, F.t1));
    F.area = rec.a[3].x;
    F.bc = rec.cr.z;
    const int p = rec.cr.x, n = rec.cr.y;
    X.p = p; X.n = n;
    Vec<double> xf = mk<double>(rec.a[3].y, rec.a[4].x, rec.a[4].y);
    Vec<double> xp = centroid(g, p), xn = centroid(g, n);
    F.l2r = vsub(xn, xp);
    S q[NV], ph[NV];
    Vec<S> gp[NV];
    ld_.load(p, false, q, gp, ph);
    recon<NV>(P, X.L, q, gp, ph, vsub(xf, xp));
    Vec<S> gn[NV];
    {
        S qn[NV], phn[NV];
        const bool bnd = F.bc >= 0;
        ld_.load(n, bnd, qn, gn, phn);
        CState<S> Rb;
        if (bnd) {
            Rb.p_center = qn[0];
            Rb.V_center = mk<S>(qn[1], qn[2], qn[3]);
            Rb.T_center = qn[4];
            if constexpr (NV == 6) Rb.nut_center = qn[5];
            else Rb.nut_center = S(0.);
        } else {
            recon<NV>(P, Rb, qn, gn, phn, vsub(xf, xn));
        }
        X.R = Rb;
    }
    if (VISC) {
        const double wp = rec.a[5].x, wn = rec.a[5].y;
        const int sl = g.wslot ? g.wslot[f] : -1;   // a wider face-average stencil's slot
        if (F.bc >= 0) {
            X.fs.gu = gp[1]; X.fs.gv = gp[2]; X.fs.gw = gp[3]; X.fs.gT = gp[4];
        } else if (sl >= 0) {
            // face average (variable.h:134-147) over a wider stencil (WTLI or
            // IDW above 2 cells), formed by the gradient pass
            X.fs.gu = ld_.face_grad(g, sl, 0);
            X.fs.gv = ld_.face_grad(g, sl, 1);
            X.fs.gw = ld_.face_grad(g, sl, 2);
            X.fs.gT = ld_.face_grad(g, sl, 3);
        } else {   // face average (variable.h:134-147), two-entry stencil in key order
            X.fs.gu = vadd(vsc(wp, gp[1]), vsc(wn, gn[1]));
            X.fs.gv = vadd(vsc(wp, gp[2]), vsc(wn, gn[2]));
            X.fs.gw = vadd(vsc(wp, gp[3]), vsc(wn, gn[3]));
            X.fs.gT = vadd(vsc(wp, gp[4]), vsc(wn, gn[4]));
        }
        if constexpr (NV == 6) {
            if (F.bc >= 0) X.fs.gn = gp[5];
            else if (sl >= 0) X.fs.gn = ld_.face_grad(g, sl, 4);
            else X.fs.gn = vadd(vsc(wp, gp[5]), vsc(wn, gn[5]));
        }
        if constexpr (VISC == kViscRans) {
            // rans.mu_t.face (ns_diffusive_face_flux.cc:37; Variable::face_calculate,
            // variable.h:134-147): the BC's value where mu_t is fixed on it
            // (inlets, no-slip walls), else the two-cell face average in key order
            const BcDev* bcp = F.bc >= 0 ? &g.bc[F.bc] : nullptr;
            if (bcp && bcp->kw_mfix) {
                X.mut_face = bcp->kw_mut;
            } else {
                double t = 0.;
                t += wp * g.kw_mut[p];
                t += wn * g.kw_mut[n];
                X.mut_face = t;
            }
        }
    }
}

// Stage 2: finish the states, apply the boundary condition (right state of
// a boundary face = a copy of the finished left state with the ghost's
// centre values, then apply_bcs; right_state_update :298-303) and evaluate
// the convective and diffusive fluxes.  PSEED (dual S only): the values of
// BC P.seed.index carry the parameter tangent's direction.
template <int FLUX, int VISC, int NV, typename S, bool PSEED = false>
__device__ __forceinline__ void face_tail(const DevGrid& g, const Params& P, int f, FaceCtx<S, NV>& X, int* status) {
    FaceGeo& F = X.F;
    finish_state(P, X.L, F);
    {
        CState<S> Rb;
        if (F.bc >= 0) {
            Rb = X.L;
            Rb.p_center = X.R.p_center;
            Rb.V_center = X.R.V_center;
            Rb.T_center = X.R.T_center;
            Rb.nut_center = X.R.nut_center;
            int flag;
            if constexpr (PSEED) {
                if (F.bc == P.seed.index) flag = apply_bcs(P, bc_val<S>(g.bc[F.bc], P.seed), X.L, Rb, F);
                else flag = apply_bcs(P, g.bc[F.bc], X.L, Rb, F);
            } else {
                flag = apply_bcs(P, g.bc[F.bc], X.L, Rb, F);
            }
            if (flag) atomicOr(status, 1);
        } else {
            Rb = X.R;
        }
        X.R = Rb;
    }
    finish_state(P, X.R, F);
    convective_flux<FLUX, NV>(P, X.L, X.R, F, FLUX == FCFD_FLUX_SW ? g.fdiss[f] : 0.0, FLUX == FCFD_FLUX_SW ? g.fwd[f] : 0.0, X.C);
    if (VISC) {
        face_state<NV>(P, X.L, X.R, F, X.fs);
        if constexpr (VISC == kViscRans) X.fs.mu_t = S(X.mut_face);
        diffusive_flux<NV, VISC == kViscRans>(P, X.fs, F, g.bc, X.D);
        if constexpr (PSEED) {   // a FIXED_Q wall's heat flux q A (ns_diffusive_face_flux.cc:58-60)
            if (F.bc >= 0 && F.bc == P.seed.index && g.bc[F.bc].thermal == FCFD_THERMAL_FIXED_Q)
                X.D[4] = X.D[4] + S(0., P.seed.qdot * F.area);
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) X.D[v] = S(0.0);
    }
}

template <int FLUX, int VISC, int NV, typename S, typename Loader, bool PSEED = false>
__device__ __forceinline__ void face_base(const DevGrid& g, const Params& P, const Loader& ld_, int f, FaceCtx<S, NV>& X, int* status) {
    face_inputs<VISC, NV>(g, ld_, P, f, X);
    face_tail<FLUX, VISC, NV, S, PSEED>(g, P, f, X, status);
}

}  // namespace fcfd
