// Device-side physics of the residual path: material model, boundary
// conditions, MUSCL face states and the convective/diffusive face fluxes.
// Templated on the scalar type S (double for the residual; a forward-mode
// dual type for the tangent kernels).  Operation order follows the reference
// so that the FP64 result rounds like it (files cited per function).
#pragma once
#include <hip/hip_runtime.h>

#include "fcfd_internal.h"

namespace fcfd {

// ---- scalar helpers ---------------------------------------------------------
__host__ __device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double sq_root(double x) { return sqrt(x); }
__device__ __forceinline__ double powr(double x, double y) { return pow(x, y); }
__device__ __forceinline__ double absv(double x) { return fabs(x); }
template <typename S> __device__ __forceinline__ S smin(const S& a, const S& b) { return (val(b) < val(a)) ? b : a; }   // std::min
template <typename S> __device__ __forceinline__ S smax(const S& a, const S& b) { return (val(a) < val(b)) ? b : a; }   // std::max

template <typename S> struct Vec { S x, y, z; };
template <typename S> __device__ __forceinline__ Vec<S> mk(S a, S b, S c) { Vec<S> r; r.x = a; r.y = b; r.z = c; return r; }
template <typename S> __device__ __forceinline__ Vec<S> vadd(const Vec<S>& a, const Vec<S>& b) { return mk<S>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename S> __device__ __forceinline__ Vec<S> vsub(const Vec<S>& a, const Vec<S>& b) { return mk<S>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename S, typename T> __device__ __forceinline__ Vec<S> vsc(const T& s, const Vec<S>& a) { return mk<S>(s * a.x, s * a.y, s * a.z); }
template <typename S, typename T> __device__ __forceinline__ S vdot(const Vec<S>& a, const Vec<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename S> __device__ __forceinline__ S vmag(const Vec<S>& a) { return sq_root(a.x * a.x + a.y * a.y + a.z * a.z); }
template <typename S> __device__ __forceinline__ Vec<S> vnorm(const Vec<S>& a) { S m = vmag(a); return mk<S>(a.x / m, a.y / m, a.z / m); }
template <typename S> __device__ __forceinline__ Vec<S> vcross(const Vec<S>& a, const Vec<S>& b) {
    return mk<S>(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
template <typename S> __device__ __forceinline__ Vec<S> lift(const Vec<double>& a) { return mk<S>(S(a.x), S(a.y), S(a.z)); }

// ---- face geometry (double only) -----------------------------------------
struct FaceGeo {
    Vec<double> n, t1, t2, l2r;
    double area;
    int bc;
};

// ---- material (src/material/material.cc:108-153) ---------------------------
template <typename S> __device__ __forceinline__ S m_rho(const Params& P, const S& p, const S& T) { return (p + P.Pref) / (P.R * (T + P.Tref)); }
template <typename S> __device__ __forceinline__ S m_p(const Params& P, const S& rho, const S& T) { return rho * P.R * (T + P.Tref) - P.Pref; }
template <typename S> __device__ __forceinline__ S m_T(const Params& P, const S& p, const S& rho) { return (p + P.Pref) / (P.R * rho) - P.Tref; }
template <typename S> __device__ __forceinline__ S m_a(const Params& P, const S& p, const S& T) { return sq_root(P.gamma * (p + P.Pref) / m_rho(P, p, T)); }
template <typename S> __device__ __forceinline__ S m_visc(const Params& P, const S& T) {
    if (P.visc_model == FCFD_VISC_SUTHERLAND)
        return P.sut_mu_ref * powr((T + P.Tref) / P.sut_T_ref, 1.5) * (P.sut_T_ref + P.sut_S) / (T + P.Tref + P.sut_S);
    return S(P.mu);
}
template <typename S> __device__ __forceinline__ S m_cond(const Params& P, const S& T) {
    if (P.cond_model == FCFD_COND_PRANDTL) return P.Cp * m_visc(P, T) / P.Pr;
    return S(P.lambda);
}

// ---- Spalart-Allmaras model (no reference counterpart) ----------------------
// SA without ft2, with the S~ modification and the negative-nu~ branch of
// Allmaras, Johnson & Spalart (ICCFD7-1902, 2012); compressible form without
// the grad(rho).grad(nu~) term:
//   d(rho nu~)/dt + div(rho V nu~) = rho (P - D)
//       + (1/sigma) [div((mu + rho nu~ fn) grad nu~) + cb2 rho |grad nu~|^2]
constexpr double sa_cb1 = 0.1355, sa_sigma = 2. / 3., sa_cb2 = 0.622, sa_kappa = 0.41, sa_cw2 = 0.3, sa_cw3 = 2.,
                 sa_cv1 = 7.1, sa_ct3 = 1.2, sa_cv2 = 0.7, sa_cv3 = 0.9, sa_cn1 = 16.;
constexpr double sa_cw1 = sa_cb1 / (sa_kappa * sa_kappa) + (1. + sa_cb2) / sa_sigma;

template <typename S> __device__ __forceinline__ S sa_fv1(const S& chi) {
    S c3 = chi * chi * chi;
    return c3 / (c3 + sa_cv1 * sa_cv1 * sa_cv1);
}
// eddy viscosity rho nu~ fv1 (nu~ > 0), else 0
template <typename S> __device__ __forceinline__ S sa_mut(const S& rho, const S& nut, const S& mu) {
    if (!(val(nut) > 0.)) return S(0.);
    S chi = rho * nut / mu;
    return rho * nut * sa_fv1(chi);
}
// diffusion factor fn (1 for nu~ >= 0)
template <typename S> __device__ __forceinline__ S sa_fn(const S& rho, const S& nut, const S& mu) {
    if (val(nut) >= 0.) return S(1.);
    S chi = rho * nut / mu;
    S c3 = chi * chi * chi;
    return (sa_cn1 + c3) / (sa_cn1 - c3);
}
// source per unit volume: rho (P - D) + (cb2/sigma) rho |grad nu~|^2, from the
// cell state, its velocity gradient (vorticity magnitude) and wall distance d
template <typename S>
__device__ __forceinline__ S sa_source(const S& rho, const S& mu, const S& nut, const Vec<S>& gu, const Vec<S>& gv, const Vec<S>& gw,
                                       const Vec<S>& gn, double d) {
    S wx = gw.y - gv.z, wy = gu.z - gw.x, wz = gv.x - gu.y;
    S Om = sq_root(wx * wx + wy * wy + wz * wz);
    S Pr, De;
    if (val(nut) >= 0.) {
        S chi = rho * nut / mu;
        S fv1 = sa_fv1(chi);
        S fv2 = 1. - chi / (1. + chi * fv1);
        double k2d2 = sa_kappa * sa_kappa * d * d;
        S Sbar = nut * fv2 / k2d2;
        S St;
        if (val(Sbar) >= val(-sa_cv2 * Om)) St = Om + Sbar;
        else St = Om + Om * (sa_cv2 * sa_cv2 * Om + sa_cv3 * Sbar) / ((sa_cv3 - 2. * sa_cv2) * Om - Sbar);
        S r = (val(St) > 0.) ? smin(nut / (St * k2d2), S(10.)) : S(10.);
        S r2 = r * r;
        S g = r + sa_cw2 * (r2 * r2 * r2 - r);
        S g2 = g * g;
        constexpr double c2 = sa_cw3 * sa_cw3;
        constexpr double c6 = c2 * c2 * c2;
        S fw = g * powr((1. + c6) / (g2 * g2 * g2 + c6), 1. / 6.);
        Pr = sa_cb1 * St * nut;
        De = sa_cw1 * fw * (nut / d) * (nut / d);
    } else {
        Pr = sa_cb1 * (1. - sa_ct3) * Om * nut;
        De = -1. * sa_cw1 * (nut / d) * (nut / d);
    }
    return rho * (Pr - De) + sa_cb2 / sa_sigma * rho * vdot(gn, gn);
}

// ---- cell/face state caches (ns_state_cache.h:26-59) -----------------------
template <typename S> struct CState {
    S p, p_center, T, T_center, rho, a, H;
    Vec<S> V, V_center, Vn;
    S nut, nut_center;   // SA nu~ (zero without the model)
};
template <typename S> struct FState {
    S T, mu, lambda;
    Vec<S> V, gu, gv, gw, gT;
    Vec<S> gn;           // SA: face gradient of nu~
    S nut, rho, mu_t;    // SA: face nu~, rho, eddy viscosity (mu_t = 0 without the model)
};

template <typename S> __device__ __forceinline__ void finish_state(const Params& P, CState<S>& s, const FaceGeo& F) {
    s.a = m_a(P, s.p, s.T);
    s.H = s.a * s.a / (P.gamma - 1.) + 0.5 * vdot(s.V, s.V);
    s.Vn = mk<S>(vdot(s.V, F.n), vdot(s.V, F.t1), vdot(s.V, F.t2));
}

// MUSCL reconstruction (ns_assemble_linear_system.cc:269-294,305-320):
// q_f = q_c + phi * (grad q . (x_f - x_c)); nu~ the same way with the SA model
template <int NV, typename S> __device__ __forceinline__ void recon(const Params& P, CState<S>& s, const S q[NV], const Vec<S> g[NV],
                                                           const S phi[NV], const Vec<double>& c2f) {
    s.p_center = q[0];
    s.p = s.p_center + phi[0] * vdot(g[0], c2f);
    s.V_center = mk<S>(q[1], q[2], q[3]);
    s.V = vadd(s.V_center, mk<S>(phi[1] * vdot(g[1], c2f), phi[2] * vdot(g[2], c2f), phi[3] * vdot(g[3], c2f)));
    s.T_center = q[4];
    s.T = s.T_center + phi[4] * vdot(g[4], c2f);
    s.rho = m_rho(P, s.p, s.T);
    if constexpr (NV == 6) {
        s.nut_center = q[5];
        s.nut = s.nut_center + phi[5] * vdot(g[5], c2f);
    } else {
        s.nut = s.nut_center = S(0.);
    }
}

// A BC's values as the scalar type S: the parameter tangent seeds them
// (BcSeed), so that the derivatives of the residual with respect to the BC
// values follow from the same apply_bcs.  The kinds, flags and the BC area
// are carried as they are.
template <typename S> struct BcVal {
    int type, kind, specified, thermal;
    S p, T, rho, V[3], p_total, T_total, mdot, nut;
    double total_area;
};
template <typename S> __device__ __forceinline__ BcVal<S> bc_val(const BcDev& b, const BcSeed& d) {
    BcVal<S> r;
    r.type = b.type; r.kind = b.kind; r.specified = b.specified; r.thermal = b.thermal;
    r.p = S(b.p, d.p); r.T = S(b.T, d.T); r.rho = S(b.rho, d.rho);
    r.V[0] = S(b.V[0], d.V[0]); r.V[1] = S(b.V[1], d.V[1]); r.V[2] = S(b.V[2], d.V[2]);
    r.p_total = S(b.p_total, d.p_total); r.T_total = S(b.T_total, d.T_total); r.mdot = S(b.mdot, d.mdot); r.nut = S(b.nut, d.nut);
    r.total_area = b.total_area;
    return r;
}

// Boundary conditions (src/navier_stokes/ns_apply_bcs.cc:25-258).  Returns a
// nonzero flag where the reference would exit (supersonic velocity inlet
// without a complete state).  The right state is carried in SSA locals and
// written back whole, so that no branch stores to an individual field of the
// state struct (divergent field stores would pin the struct in scratch).
// B: BcDev (the BC's values as constants) or BcVal<S> (seeded values).
template <typename S, typename B> __device__ __forceinline__ int apply_bcs(const Params& P, const B& bc, const CState<S>& L,
                                                      CState<S>& Rout, const FaceGeo& F) {
    const double g = P.gamma, Pref = P.Pref, Tref = P.Tref;
    int flag = 0;
    const Vec<S> n = lift<S>(F.n);
    S p = Rout.p, pc = Rout.p_center, T = Rout.T, Tc = Rout.T_center, rho = Rout.rho, a = Rout.a;
    Vec<S> V = Rout.V, Vc = Rout.V_center;
    S nt = Rout.nut, ntc = Rout.nut_center;
    if (bc.type == FCFD_BC_INLET) {
        if (bc.kind == FCFD_KIND_VELOCITY) {
            V = mk<S>(S(bc.V[0]), S(bc.V[1]), S(bc.V[2]));
            S uNL = vdot(L.V, F.n);
            S uNR = vdot(V, F.n);
            a = L.a + 0.5 * (uNL - uNR) * (g - 1.);
            S MachL = -uNL / L.a;
            if (val(MachL) >= 1. && bc.specified != FCFD_SPEC_STATE) flag = 1;
            if (bc.specified == FCFD_SPEC_STATE) { p = S(bc.p); T = S(bc.T); rho = S(bc.rho); }
            else if (bc.specified == FCFD_SPEC_P) { p = S(bc.p); rho = g * (p + Pref) / (a * a); T = m_T(P, p, rho); }
            else if (bc.specified == FCFD_SPEC_T) { T = S(bc.T); rho = L.rho * powr((L.T + Tref) / (T + Tref), 1. / (1. - g)); p = m_p(P, rho, T); }
            else { rho = S(bc.rho); p = a * a * rho / g - Pref; T = m_T(P, p, rho); }
            pc = p; Vc = V; Tc = T;
        } else if (bc.kind == FCFD_KIND_MDOT) {
            S uNL = vdot(L.V, F.n);
            const auto mdotNR = -1. * bc.mdot / bc.total_area;
            // the reference's test `!specified != BC_STATE` (ns_apply_bcs.cc:99-104)
            // is true for every specification: a supersonic mass-flow inlet always stops
            if (val(-uNL / L.a) >= 1.) flag = 1;
            if (bc.specified == FCFD_SPEC_STATE) { p = S(bc.p); T = S(bc.T); rho = S(bc.rho); }
            else if (bc.specified == FCFD_SPEC_P) { p = S(bc.p); T = L.T; rho = m_rho(P, p, T); }
            else if (bc.specified == FCFD_SPEC_T) { T = S(bc.T); p = L.p; rho = m_rho(P, p, T); }
            else {
                rho = S(bc.rho);
                a = 0.5 * (g - 1.) * (uNL + 2. * L.a / (g - 1.) - mdotNR / rho);
                p = a * a * rho / g - Pref;
                T = m_T(P, p, rho);
            }
            V = vsc(mdotNR / rho, n);
            pc = p; Vc = V; Tc = T;
        } else if (bc.kind == FCFD_KIND_STAGNATION) {
            if (val(vdot(L.V, F.n)) > 0.) V = vsc(-1., L.V_center);
            else V = L.V_center;
            T = bc.T_total - 0.5 * (g - 1.) / (g * P.R) * vdot(V, V);
            p = (bc.p_total + Pref) / (1. + 0.5 * vdot(V, V) / (P.R * (T + Tref))) - Pref;
            rho = m_rho(P, p, T);
            pc = p; Vc = V; Tc = T;
        }
    } else if (bc.type == FCFD_BC_OUTLET) {
        S uNL = vdot(L.V, F.n);
        S MachL = uNL / L.a;
        bool supersonic = val(MachL) > 1.;
        if (bc.specified == FCFD_SPEC_P && !supersonic) {
            p = S(bc.p); pc = p;
            rho = L.rho * powr((p + Pref), 1. / g) / powr((L.p + Pref), 1. / g);
            a = sq_root(g * (p + Pref) / rho);
            T = m_T(P, p, rho);
            V = L.V;
        } else if (bc.specified == FCFD_SPEC_T && !supersonic) {
            T = S(bc.T); Tc = T;
            a = sq_root(g * P.R * (T + Tref));
            rho = powr(g * (L.p + Pref) / (a * a * powr(L.rho, g)), 1. - g);
            p = m_p(P, rho, T);
            V = L.V;
        } else if (bc.specified == FCFD_SPEC_RHO && !supersonic) {
            rho = S(bc.rho);
            p = (L.p + Pref) * powr(rho / L.rho, g) - Pref;
            T = m_T(P, p, rho);
            a = sq_root(g * (p + Pref) / rho);
            V = L.V;
        } else {
            // extrapolate everything; the reference leaves rho stale here, the
            // density of the extrapolated p, T is used instead
            p = L.p; T = L.T; V = L.V;
            rho = m_rho(P, p, T);
        }
        if (val(MachL) < 0.) {
            if (bc.kind == FCFD_KIND_DAMP_REVERSE) {
                p = p - 0.5 * L.rho * uNL * uNL;
                rho = L.rho * powr((p + Pref), 1. / g) / powr((L.p + Pref), 1. / g);
                a = sq_root(g * (p + Pref) / rho);
                T = m_T(P, p, rho);
            }
            if (bc.kind == FCFD_KIND_NO_REVERSE) V = vsc(-1., L.V);
        }
        if (val(MachL) < 1. && bc.kind == FCFD_KIND_FORCE_SUPERSONIC) {
            S d = L.a - uNL;
            p = p - 0.5 * L.rho * (d * d);
            p = smin(S(1.e-8), p + Pref) - Pref;
            rho = L.rho * powr((p + Pref), 1. / g) / powr((L.p + Pref), 1. / g);
            a = sq_root(g * (p + Pref) / rho);
            T = m_T(P, p, rho);
        }
    } else if (bc.type == FCFD_BC_WALL) {
        if (bc.thermal == FCFD_THERMAL_FIXED_T) {
            T = S(bc.T); Tc = T;
            p = L.p;
            rho = m_rho(P, p, T);
        } else {
            p = L.p; rho = L.rho; T = L.T; Tc = L.T_center;
        }
        if (bc.kind == FCFD_KIND_SLIP) {
            V = vsub(L.V, vsc(2. * vdot(L.V, F.n), n));
            Vc = vsub(L.V_center, vsc(2. * vdot(L.V_center, F.n), n));
        } else {
            V = vsc(-1., L.V);
            Vc = vsc(-1., L.V_center);
        }
        pc = L.p_center;
    } else if (bc.type == FCFD_BC_SYMMETRY) {
        p = L.p;
        V = vsub(L.V, vsc(2. * vdot(L.V, F.n), n));
        T = L.T;
        rho = L.rho;
        pc = L.p_center;
        Tc = L.T_center;
        Vc = vsub(L.V_center, vsc(2. * vdot(L.V_center, F.n), n));
    }
    // SA nu~: inlet value; zero on a no-slip wall (the ghost mirrors to -nu~);
    // copied otherwise (an outlet keeps the ghost's centre value)
    if (bc.type == FCFD_BC_INLET) { nt = S(bc.nut); ntc = S(bc.nut); }
    else if (bc.type == FCFD_BC_WALL && bc.kind != FCFD_KIND_SLIP) { nt = -1. * L.nut; ntc = -1. * L.nut_center; }
    else if (bc.type == FCFD_BC_OUTLET) { nt = L.nut; }
    else { nt = L.nut; ntc = L.nut_center; }
    Rout.p = p; Rout.p_center = pc; Rout.T = T; Rout.T_center = Tc; Rout.rho = rho; Rout.a = a;
    Rout.V = V; Rout.V_center = Vc;
    Rout.nut = nt; Rout.nut_center = ntc;
    return flag;
}

// ---- convective fluxes ----------------------------------------------------
// Roe with entropy fix (ns_roe.cc:25-98)
template <typename S> __device__ __forceinline__ void roe_flux(const CState<S>& L, const CState<S>& R, S fn[5], double G) {
    S rho = sq_root(R.rho / L.rho);
    S u = (L.Vn.x + rho * R.Vn.x) / (1. + rho);
    S v = (L.Vn.y + rho * R.Vn.y) / (1. + rho);
    S w = (L.Vn.z + rho * R.Vn.z) / (1. + rho);
    S H = (L.H + rho * R.H) / (1. + rho);
    S a = sq_root((G - 1.) * (H - 0.5 * (u * u + v * v + w * w)));
    rho = rho * L.rho;
    S Du = R.Vn.x - L.Vn.x;
    S Dp = R.p - L.p;
    if (val(u) >= 0.) {
        S deltaV = 0.5 * (Dp - rho * a * Du) / (a * a);
        S lambda = u - a;
        S Dl = 2. * (smin(a, smax(S(0.), 2. * (L.a - R.a + Du))));
        if (val(lambda) <= val(-0.5 * Dl)) {
        } else if (val(lambda) < val(0.5 * Dl)) {
            lambda = -0.5 * (lambda - 0.5 * Dl) * (lambda - 0.5 * Dl) / Dl;
        } else {
            lambda = S(0.);
        }
        S mdot = L.rho * L.Vn.x;
        S pr = lambda * deltaV;
        fn[0] = mdot + pr;
        fn[1] = mdot * L.Vn.x + L.p + pr * (u - a);
        fn[2] = mdot * L.Vn.y + pr * v;
        fn[3] = mdot * L.Vn.z + pr * w;
        fn[4] = mdot * L.H + pr * (H - u * a);
    } else {
        S deltaV = 0.5 * (Dp + rho * a * Du) / (a * a);
        S lambda = u + a;
        S Dl = 2. * (smin(a, smax(S(0.), 2. * (R.a - L.a + Du))));
        if (val(lambda) >= val(0.5 * Dl)) {
        } else if (val(lambda) > val(-0.5 * Dl)) {
            lambda = 0.5 * (lambda + 0.5 * Dl) * (lambda + 0.5 * Dl) / Dl;
        } else {
            lambda = S(0.);
        }
        S mdot = R.rho * R.Vn.x;
        S pr = lambda * deltaV;
        fn[0] = mdot - pr;
        fn[1] = mdot * R.Vn.x + R.p - pr * (u + a);
        fn[2] = mdot * R.Vn.y - pr * v;
        fn[3] = mdot * R.Vn.z - pr * w;
        fn[4] = mdot * R.H - pr * (H + u * a);
    }
}

// AUSM+-up (ns_ausm_plus_up.cc:35-152)
template <typename S> __device__ __forceinline__ S ms2p(const S& M) { return 0.25 * (M + 1.) * (M + 1.); }
template <typename S> __device__ __forceinline__ S ms2m(const S& M) { return -0.25 * (M - 1.) * (M - 1.); }
template <typename S> __device__ __forceinline__ S ms4p(const S& M) { return (val(absv(M)) >= 1.) ? S(0.5 * (M + absv(M))) : S(ms2p(M) * (1. - 16. * 0.125 * ms2m(M))); }
template <typename S> __device__ __forceinline__ S ms4m(const S& M) { return (val(absv(M)) >= 1.) ? S(0.5 * (M - absv(M))) : S(ms2m(M) * (1. + 16. * 0.125 * ms2p(M))); }
template <typename S> __device__ __forceinline__ S ps5p(const S& M, const S& al) { return (val(absv(M)) >= 1.) ? S(0.5 * (M + absv(M)) / M) : S(ms2p(M) * ((2. - M) - 16. * al * M * ms2m(M))); }
template <typename S> __device__ __forceinline__ S ps5m(const S& M, const S& al) { return (val(absv(M)) >= 1.) ? S(0.5 * (M - absv(M)) / M) : S(ms2m(M) * ((-2. - M) + 16. * al * M * ms2p(M))); }

template <typename S> __device__ __forceinline__ void ausm_flux(const CState<S>& L, const CState<S>& R, S fn[5], double Pref, double Minf) {
    const double Kp = 0.25, Ku = 0.75, sigma = 1.;
    S aLh = L.a * L.a / smax(L.a, L.Vn.x);
    S aRh = R.a * R.a / smax(R.a, -1. * R.Vn.x);
    S a = smin(aLh, aRh);
    S rho = 0.5 * (L.rho + R.rho);
    S ML = L.Vn.x / a, MR = R.Vn.x / a;
    S Mbar2 = 0.5 * (ML * ML + MR * MR);
    S Mref = smax(S(Minf), sq_root(Mbar2));
    Mref = smin(Mref, S(1.));
    S fa;
    if (val(Mbar2) >= 1.) fa = S(1.);
    else { S Mo = sq_root(smin(S(1.), smax(Mbar2, Mref * Mref))); fa = Mo * (2. - Mo); }
    S M = ms4p(ML) + ms4m(MR) - Kp / fa * smax(1. - sigma * Mbar2, S(0.)) * (R.p - L.p) / (rho * a * a);
    S mdot = (val(M) > 0) ? S(a * M * L.rho) : S(a * M * R.rho);
    S alpha = 3. / 16. * (-4. + 5. * fa * fa);
    S pp = ps5p(ML, alpha), pm = ps5m(MR, alpha);
    S p = pp * (L.p + Pref) + pm * (R.p + Pref) - Ku * pp * pm * (L.rho + R.rho) * fa * a * (R.Vn.x - L.Vn.x);
    p = p - Pref;
    fn[0] = mdot;
    if (val(mdot) > 0.) { fn[1] = mdot * L.Vn.x + p; fn[2] = mdot * L.Vn.y; fn[3] = mdot * L.Vn.z; fn[4] = mdot * L.H; }
    else { fn[1] = mdot * R.Vn.x + p; fn[2] = mdot * R.Vn.y; fn[3] = mdot * R.Vn.z; fn[4] = mdot * R.H; }
}

// van Leer (ns_vanleer.cc:25-66)
template <typename S> __device__ __forceinline__ void vanleer_flux(const CState<S>& L, const CState<S>& R, S fn[5], double G, double Pref) {
    S ML = L.Vn.x / L.a, MR = R.Vn.x / R.a;
    S M = 0.5 * (ML + MR);
    if (val(M) <= -1.) {
        S mdot = R.rho * R.Vn.x;
        fn[0] = mdot; fn[1] = mdot * R.Vn.x + R.p; fn[2] = mdot * R.Vn.y; fn[3] = mdot * R.Vn.z; fn[4] = mdot * R.H;
    } else if (val(M) >= 1.) {
        S mdot = L.rho * L.Vn.x;
        fn[0] = mdot; fn[1] = mdot * L.Vn.x + L.p; fn[2] = mdot * L.Vn.y; fn[3] = mdot * L.Vn.z; fn[4] = mdot * L.H;
    } else {
        S fp = 0.25 * L.rho * L.a * (ML + 1.) * (ML + 1.);
        S fm = -0.25 * R.rho * R.a * (1. - MR) * (1. - MR);
        fn[0] = fp + fm;
        S tL = (2. * L.a / G) * (0.5 * (G - 1) * ML + 1.);
        S tR = (2. * R.a / G) * (0.5 * (G - 1) * MR - 1.);
        fn[1] = fp * tL + fm * tR - Pref;
        fn[2] = fp * L.Vn.y + fm * R.Vn.y;
        fn[3] = fp * L.Vn.z + fm * R.Vn.z;
        tL = 0.5 * G * G / (G * G - 1.) * tL * tL + 0.5 * (L.Vn.y * L.Vn.y + L.Vn.z * L.Vn.z);
        tR = 0.5 * G * G / (G * G - 1.) * tR * tR + 0.5 * (R.Vn.y * R.Vn.y + R.Vn.z * R.Vn.z);
        fn[4] = fp * tL + fm * tR;
    }
}

// SD-SLAU (ns_sd_slau.cc:37-94)
template <typename S> __device__ __forceinline__ S sgn(const S& x) { return (val(x) < 0) ? S(-1.) : (val(x) > 0) ? S(1.) : S(0.); }
template <typename S> __device__ __forceinline__ void slau_flux(const CState<S>& L, const CState<S>& R, S fn[5], double Pref) {
    const double Csd1 = 0.1, Csd2 = 10.;
    S a = 0.5 * (L.a + R.a);
    S Mhat = smin(S(1.), sq_root(0.5 * (vdot(L.V, L.V) + vdot(R.V, R.V)) / a));
    S chi = (1. - Mhat) * (1. - Mhat);
    S Mp = L.Vn.x / a, Mm = R.Vn.x / a;
    const double alpha = 0.;
    S Bp, Bm;
    if (val(absv(Mp)) < 1) Bp = 0.25 * (2. - Mp) * (Mp + 1.) * (Mp + 1.) + alpha * Mp * (Mp - 1.) * (Mp - 1.);
    else Bp = 0.5 * (1. + sgn(Mp));
    if (val(absv(Mm)) < 1) Bm = 0.25 * (2. + Mm) * (Mm - 1.) * (Mm - 1.) - alpha * Mm * (Mm - 1.) * (Mm - 1.);
    else Bm = 0.5 * (1. + sgn(S(-1. * Mm)));
    S ave_p = 0.5 * (L.p + R.p) + Pref;
    S dp = R.p - L.p;
    S p = ave_p - 0.5 * (Bp - Bm) * dp + (1. - chi) * (Bp + Bm - 1.) * ave_p;
    S maxdp = absv(dp);
    S gg = -1. * smax(smin(Mp, S(0.)), S(-1.)) * smin(smax(Mm, S(0.)), S(1.));
    S Vn = (L.rho * absv(L.Vn.x) + R.rho * absv(R.Vn.x)) / (L.rho + R.rho);
    S Vnp = (1. - gg) * Vn + gg * absv(L.Vn.x);
    S Vnm = (1. - gg) * Vn + gg * absv(R.Vn.x);
    S th = (Csd2 * absv(dp) / ave_p + Csd1) / (maxdp / ave_p + Csd1);
    th = smin(S(1.), th * th);
    S mdot = 0.5 * (L.rho * (L.Vn.x + Vnp) + R.rho * (R.Vn.x - Vnm) - th * smax(S(0.), (1. - Vn / a)) * dp / a);
    p = p - Pref;
    fn[0] = mdot;
    if (val(mdot) > 0.) { fn[1] = mdot * L.Vn.x + p; fn[2] = mdot * L.Vn.y; fn[3] = mdot * L.Vn.z; fn[4] = mdot * L.H; }
    else { fn[1] = mdot * R.Vn.x + p; fn[2] = mdot * R.Vn.y; fn[3] = mdot * R.Vn.z; fn[4] = mdot * R.H; }
}

// modified Steger-Warming (ns_stegger_warming.cc:23-145)
template <typename S> __device__ __forceinline__ void sw_flux(const Params& P, const CState<S>& L, const CState<S>& R, double diss,
                                                     double cwd, S fn[5]) {
    const double alpha = 6.;
    const double RW = P.R, gm = P.gamma;
    for (int i = 0; i < 5; ++i) fn[i] = S(0.);
    S deltap = absv(L.p - R.p) / smin(L.p, R.p);
    S w = 0.5 / (alpha * alpha * deltap * deltap + 1.);
    for (int h = 0; h <= 1; h++) {
        double signal = h ? -1. : 1.;
        S wL = (signal > 0.) ? S(1. - w) : w;
        S wR = 1. - wL;
        S p = wL * L.p + wR * R.p;
        Vec<S> V = vadd(vsc(wL, L.Vn), vsc(wR, R.Vn));
        S V2 = vmag(V); V2 = V2 * V2;
        S T = wL * L.T + wR * R.T;
        S rho = wL * L.rho + wR * R.rho;
        double Cp = P.Cp;
        double beta = gm * RW / Cp;
        S a = sq_root((1. + beta) * p / rho);
        S a2 = a * a;
        S H = 0.5 * V2 + a2 / (gm - 1.);
        S gs = RW * T + beta * (0.5 * V2 - Cp * T / gm);
        S eta = (beta * V2 - gs) / (beta * a2);
        S Lm[5][5], Rm[5][5], Lam[5];
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) { Lm[i][j] = S(0.); Rm[i][j] = S(0.); }
        Lm[0][0] = 1.0 / a2;
        Lm[0][3] = Lm[0][4] = 0.5 / a2;
        Lm[1][0] = V.x / a2; Lm[2][0] = V.y / a2; Lm[3][0] = V.z / a2;
        Lm[1][3] = 0.5 * (V.x + a) / a2;
        Lm[1][4] = 0.5 * (V.x - a) / a2;
        Lm[2][1] = S(1.);
        Lm[2][3] = Lm[2][4] = 0.5 * V.y / a2;
        Lm[3][2] = S(1.);
        Lm[3][3] = Lm[3][4] = 0.5 * V.z / a2;
        Lm[4][0] = eta; Lm[4][1] = V.y; Lm[4][2] = V.z;
        Lm[4][3] = 0.5 * (H + a * V.x) / a2;
        Lm[4][4] = 0.5 * (H - a * V.x) / a2;
        S eps = (cwd > P.bl_height) ? S(P.wdiss * (a + vmag(V))) : S(P.wdiss * diss * (a + absv(V.x)));
        Lam[0] = Lam[1] = Lam[2] = 0.5 * (V.x + signal * sq_root(V.x * V.x + eps * eps));
        Lam[3] = 0.5 * (V.x + a + signal * sq_root((V.x + a) * (V.x + a) + eps * eps));
        Lam[4] = 0.5 * (V.x - a + signal * sq_root((V.x - a) * (V.x - a) + eps * eps));
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Rm[i][j] = Rm[i][j] + Lm[i][j] * Lam[j];
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Lm[i][j] = Rm[i][j];
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Rm[i][j] = S(0.);
        Rm[0][0] = a2 - gs; Rm[0][1] = beta * V.x; Rm[0][2] = beta * V.y; Rm[0][3] = beta * V.z; Rm[0][4] = S(-beta);
        Rm[1][0] = -1. * V.y; Rm[1][2] = S(1.);
        Rm[2][0] = -1. * V.z; Rm[2][3] = S(1.);
        Rm[3][0] = gs - V.x * a; Rm[3][1] = a - beta * V.x; Rm[3][2] = -beta * V.y; Rm[3][3] = -beta * V.z;
        Rm[4][0] = gs + V.x * a; Rm[4][1] = -1. * a - beta * V.x; Rm[4][2] = -beta * V.y; Rm[4][3] = -beta * V.z;
        Rm[3][4] = Rm[4][4] = S(beta);
        S Q[5];
        if (signal > 0) { Q[0] = L.rho; Q[1] = L.rho * L.Vn.x; Q[2] = L.rho * L.Vn.y; Q[3] = L.rho * L.Vn.z; Q[4] = L.rho * L.H - L.p; }
        else { Q[0] = R.rho; Q[1] = R.rho * R.Vn.x; Q[2] = R.rho * R.Vn.y; Q[3] = R.rho * R.Vn.z; Q[4] = R.rho * R.H - R.p; }
        for (int i = 0; i < 5; ++i) {
            for (int j = 0; j < 5; ++j) {
                S J = S(0.);
                for (int k = 0; k < 5; k++) J = J + Lm[i][k] * Rm[k][j];
                fn[i] = fn[i] + J * Q[j];
            }
        }
    }
}

// weightL.face as each flux function leaves it (ns_roe.cc:95,
// ns_ausm_plus_up.cc:93,99, ns_sd_slau.cc:84,90, ns_vanleer.cc:39,48,62,
// ns_stegger_warming.cc:38-46): Roe 0.5; van Leer 0 / 1 / 0.5 by the mean
// normal Mach number; Steger-Warming the pressure-switch weight of its second
// pass; AUSM+-up and SD-SLAU 1 where the face mass flux c0 is positive, else 0
template <int FLUX>
__device__ __forceinline__ double weight_left(const CState<double>& L, const CState<double>& R, double c0) {
    if (FLUX == FCFD_FLUX_ROE) return 0.5;
    if (FLUX == FCFD_FLUX_VAN_LEER) {
        const double M = 0.5 * (L.Vn.x / L.a + R.Vn.x / R.a);
        return (M <= -1.) ? 0. : (M >= 1.) ? 1. : 0.5;
    }
    if (FLUX == FCFD_FLUX_SW) {
        const double deltap = fabs(L.p - R.p) / smin(L.p, R.p);
        return 0.5 / (6. * 6. * deltap * deltap + 1.);
    }
    return (c0 > 0.) ? 1. : 0.;
}

// dispatch + rotation back to xyz (ns_convective_face_flux.cc:32-59)
template <int FLUX, int NV, typename S>
__device__ __forceinline__ void convective_flux(const Params& P, const CState<S>& L, const CState<S>& R, const FaceGeo& F, double diss,
                                                double cwd, S flux[NV]) {
    S fn[5];
    if (FLUX == FCFD_FLUX_ROE) roe_flux(L, R, fn, P.gamma);
    else if (FLUX == FCFD_FLUX_VAN_LEER) vanleer_flux(L, R, fn, P.gamma, P.Pref);
    else if (FLUX == FCFD_FLUX_SD_SLAU) slau_flux(L, R, fn, P.Pref);
    else if (FLUX == FCFD_FLUX_SW) sw_flux(P, L, R, diss, cwd, fn);
    else ausm_flux(L, R, fn, P.Pref, P.Minf);
    flux[0] = fn[0] * F.area;
    flux[1] = (fn[1] * F.n.x + fn[2] * F.t1.x + fn[3] * F.t2.x) * F.area;
    flux[2] = (fn[1] * F.n.y + fn[2] * F.t1.y + fn[3] * F.t2.y) * F.area;
    flux[3] = (fn[1] * F.n.z + fn[2] * F.t1.z + fn[3] * F.t2.z) * F.area;
    flux[4] = fn[4] * F.area;
    if constexpr (NV == 6) flux[5] = flux[0] * ((val(flux[0]) >= 0.) ? L.nut : R.nut);   // SA: nu~ upwind with the mass flux
}

// SA face quantities: nu~ and rho averages of the two face states, mu_t
template <typename S> __device__ __forceinline__ void sa_face(const CState<S>& L, const CState<S>& R, FState<S>& fs) {
    fs.nut = 0.5 * (L.nut + R.nut);
    fs.rho = 0.5 * (L.rho + R.rho);
    fs.mu_t = sa_mut(fs.rho, fs.nut, fs.mu);
}

// deferred correction of one face gradient along l2r (ns_assemble_linear_system.cc:394-404)
template <typename S> __device__ __forceinline__ Vec<S> correct(const Vec<S>& g, const Vec<double>& l2rn, double l2rmag, const S& dq) {
    Vec<S> r = vsub(g, vsc(vdot(g, l2rn), lift<S>(l2rn)));
    return vadd(r, vsc(dq / l2rmag, lift<S>(l2rn)));
}

// face_state_update (ns_assemble_linear_system.cc:353-415), given the face
// gradients before correction
template <int NV, typename S>
__device__ __forceinline__ void face_state(const Params& P, const CState<S>& L, const CState<S>& R, const FaceGeo& F, FState<S>& fs) {
    Vec<double> l2rn = vnorm(F.l2r);
    double l2rmag = vmag(F.l2r);
    fs.gu = correct(fs.gu, l2rn, l2rmag, S(R.V_center.x - L.V_center.x));
    fs.gv = correct(fs.gv, l2rn, l2rmag, S(R.V_center.y - L.V_center.y));
    fs.gw = correct(fs.gw, l2rn, l2rmag, S(R.V_center.z - L.V_center.z));
    fs.gT = correct(fs.gT, l2rn, l2rmag, S(R.T_center - L.T_center));
    fs.V = vsc(0.5, vadd(L.V, R.V));
    fs.T = 0.5 * (L.T + R.T);
    fs.mu = m_visc(P, fs.T);
    fs.lambda = m_cond(P, fs.T);
    if constexpr (NV == 6) {
        fs.gn = correct(fs.gn, l2rn, l2rmag, S(R.nut_center - L.nut_center));
        sa_face(L, R, fs);
    } else {
        fs.mu_t = S(0.);
    }
}

// diffusive_face_flux (ns_diffusive_face_flux.cc:28-64): mu + turb_visc and
// lambda + Cp turb_visc / Pr_t (:36-39), turb_visc = mu_t(face) of the SA
// model or (RANS) of the segregated two-equation model, 0 without either;
// plus the SA diffusion of nu~
template <int NV, bool RANS = false, typename S>
__device__ __forceinline__ void diffusive_flux(const Params& P, const FState<S>& f, const FaceGeo& F, const BcDev* bc, S flux[NV]) {
    Vec<double> aV = mk<double>(F.n.x * F.area, F.n.y * F.area, F.n.z * F.area);
    Vec<S> tx, ty, tz;
    tx.x = 2. / 3. * (2. * f.gu.x - f.gv.y - f.gw.z);
    tx.y = f.gu.y + f.gv.x;
    tx.z = f.gu.z + f.gw.x;
    ty.x = tx.y;
    ty.y = 2. / 3. * (2. * f.gv.y - f.gu.x - f.gw.z);
    ty.z = f.gv.z + f.gw.y;
    tz.x = tx.z;
    tz.y = ty.z;
    tz.z = 2. / 3. * (2. * f.gw.z - f.gu.x - f.gv.y);
    S mu = f.mu + f.mu_t;
    flux[0] = S(0.);
    flux[1] = mu * vdot(tx, aV);
    flux[2] = mu * vdot(ty, aV);
    flux[3] = mu * vdot(tz, aV);
    flux[4] = mu * (vdot(tx, f.V) * aV.x + vdot(ty, f.V) * aV.y + vdot(tz, f.V) * aV.z);
    S tc = S(0.);
    if constexpr (NV == 6 || RANS) tc = P.Cp * f.mu_t / P.Pr_t;
    S qq = (f.lambda + tc) * vdot(f.gT, aV);
    if (F.bc >= 0 && bc[F.bc].thermal == FCFD_THERMAL_FIXED_Q) qq = S(bc[F.bc].qdot * F.area);
    flux[4] = flux[4] + qq;
    if constexpr (NV == 6) flux[5] = (f.mu + f.rho * f.nut * sa_fn(f.rho, f.nut, f.mu)) * vdot(f.gn, aV) / sa_sigma;
}

}  // namespace fcfd
