// Reference-equivalent finite-difference Jacobian (FP64) and its assembly.
//
//   fdjac_kernel     one thread per face: the 1 + 2 NV flux evaluations of
//                    get_jacobians (ns_assemble_linear_system.cc:197-267) with
//                    the same epsilon policy, BC re-application on left
//                    perturbations, face-state adjustment (including the
//                    case-1 fall-through and the face-state drift between
//                    perturbed variables), writing JL[var][eq] and JR[var][eq]
//                    per face ([NV*NV][F] each, coalesced).
//   diag_kernel      one thread per cell: diagonal block = sum over the cell's
//                    faces in reference face order of -JL (parent side) or
//                    +JR (neighbour side), then (SA model) the source block by
//                    one-sided differences in the cell state, then + P V / dt
//                    (time_terms, ns_time_terms.cc:147-163) -- the
//                    MatSetValues(ADD) order.
//   The off-diagonal block (row P, col N) is -JR^T and (row N, col P) is +JL^T
//   of the face; each is touched by exactly one face, so it stays per face.
// With the SA model (NV = 6, no reference counterpart) nu~ is perturbed like
// T (face and centre value), the SA face quantities follow every perturbation,
// and the nu~ step floor is 10 sqrt(eps) times the largest inlet nu~.
#include <hip/hip_runtime.h>

#include "fcfd_face.h"
#include "fcfd_jacobian.h"

namespace fcfd {

// state_perturb (:417-458)
__device__ __forceinline__ void state_perturb(const Params& P, CState<double>& s, const FaceGeo& F, int var, double eps) {
    switch (var) {
    case 0: s.p += eps; s.rho = m_rho(P, s.p, s.T); s.a = m_a(P, s.p, s.T); break;
    case 1: s.V.x += eps; s.V_center.x += eps; s.Vn = mk<double>(vdot(s.V, F.n), vdot(s.V, F.t1), vdot(s.V, F.t2)); break;
    case 2: s.V.y += eps; s.V_center.y += eps; s.Vn = mk<double>(vdot(s.V, F.n), vdot(s.V, F.t1), vdot(s.V, F.t2)); break;
    case 3: s.V.z += eps; s.V_center.z += eps; s.Vn = mk<double>(vdot(s.V, F.n), vdot(s.V, F.t1), vdot(s.V, F.t2)); break;
    case 4: s.T += eps; s.T_center += eps; s.rho = m_rho(P, s.p, s.T); s.a = m_a(P, s.p, s.T); break;
    default: s.nut += eps; s.nut_center += eps; break;
    }
    s.H = s.a * s.a / (P.gamma - 1.) + 0.5 * vdot(s.V, s.V);
}

// face_state_adjust (:460-500); case 1 falls through into case 2
template <int NV>
__device__ __forceinline__ void face_state_adjust(const Params& P, const CState<double>& L, const CState<double>& R, const FaceGeo& F,
                                                  FState<double>& f, int var) {
    Vec<double> n = vnorm(F.l2r);
    double m = vmag(F.l2r);
    if (var == 1) {
        f.V.x = 0.5 * (L.V.x + R.V.x);
        f.gu = correct(f.gu, n, m, R.V_center.x - L.V_center.x);
    }
    if (var == 1 || var == 2) {
        f.V.y = 0.5 * (L.V.y + R.V.y);
        f.gv = correct(f.gv, n, m, R.V_center.y - L.V_center.y);
    } else if (var == 3) {
        f.V.z = 0.5 * (L.V.z + R.V.z);
        f.gw = correct(f.gw, n, m, R.V_center.z - L.V_center.z);
    } else if (var == 4) {
        f.T = 0.5 * (L.T + R.T);
        f.gT = correct(f.gT, n, m, R.T_center - L.T_center);
        f.mu = m_visc(P, f.T);
        f.lambda = m_cond(P, f.T);
    }
    if constexpr (NV == 6) {   // SA face quantities follow the perturbed states
        if (var == 5) f.gn = correct(f.gn, n, m, R.nut_center - L.nut_center);
        sa_face(L, R, f);
    }
}

__device__ __forceinline__ double fd_eps(double upd, double small_number) {   // (:48, :217-219)
    const double factor = 0.001;
    if (upd > small_number) return smax(small_number, factor * upd);
    if (upd < -small_number) return smin(-small_number, factor * upd);
    return small_number;
}
__device__ __forceinline__ double fd_floor(const Params& P, int var) {
    return var == 5 ? P.nut_eps : 10. * 1.4901161193847656e-08;   // 10*sqrt(DBL_EPSILON)
}

template <int FLUX, int VISC, int NV>
__global__ void __launch_bounds__(128) fdjac_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ G,
                                                    const double* __restrict__ PHI, int has_phi, double phi_const,
                                                    const double* __restrict__ UPD, double* __restrict__ JL,
                                                    double* __restrict__ JR, int* __restrict__ status) {
    const int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    FaceCtx<double, NV> X;
    CellLoader<NV> ld_{W, G, PHI, has_phi != 0, phi_const, (size_t)g.ld};
    face_base<FLUX, VISC, NV>(g, P, ld_, f, X, status);
    const FaceGeo& F = X.F;
    const size_t ld = g.ld;
    const size_t Fs = g.F;
    const bool interior = F.bc < 0;
    const double diss = FLUX == FCFD_FLUX_SW ? g.fdiss[f] : 0.0, cwd = FLUX == FCFD_FLUX_SW ? g.fwd[f] : 0.0;
    for (int var = 0; var < NV; ++var) {
        double Cp[NV], Dp[NV];
        CState<double> Lp = X.L, Rp = X.R;
        const double sn = fd_floor(P, var);
        double eps = fd_eps(UPD[var * ld + X.p], sn);
        state_perturb(P, Lp, F, var, eps);
        if (!interior) {   // right_state_update(leftPlus, rightPlus) (:224)
            Rp.p_center = W[X.n];
            Rp.V_center = mk<double>(W[ld + X.n], W[2 * ld + X.n], W[3 * ld + X.n]);
            Rp.T_center = W[4 * ld + X.n];
            if constexpr (NV == 6) Rp.nut_center = W[5 * ld + X.n];
            if (apply_bcs(P, g.bc[F.bc], Lp, Rp, F)) atomicOr(status, 1);
            finish_state(P, Rp, F);
        }
        convective_flux<FLUX, NV>(P, Lp, Rp, F, diss, cwd, Cp);
        if (VISC) {
            face_state_adjust<NV>(P, Lp, Rp, F, X.fs, var);
            diffusive_flux<NV, VISC == kViscRans>(P, X.fs, F, g.bc, Dp);
        } else {
#pragma unroll
            for (int j = 0; j < NV; ++j) Dp[j] = 0.0;
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) JL[pidx(Fs, NV * var + j, f)] = (Dp[j] - X.D[j] - Cp[j] + X.C[j]) / eps;
        if (interior) {
            Rp = X.R;
            double eps2 = fd_eps(UPD[var * ld + X.n], sn);
            state_perturb(P, Rp, F, var, eps2);
            convective_flux<FLUX, NV>(P, X.L, Rp, F, diss, cwd, Cp);
            if (VISC) {
                face_state_adjust<NV>(P, X.L, Rp, F, X.fs, var);
                diffusive_flux<NV, VISC == kViscRans>(P, X.fs, F, g.bc, Dp);
            }
#pragma unroll
            for (int j = 0; j < NV; ++j) JR[pidx(Fs, NV * var + j, f)] = (Dp[j] - X.D[j] - Cp[j] + X.C[j]) / eps2;
        } else {
#pragma unroll
            for (int j = 0; j < NV; ++j) JR[pidx(Fs, NV * var + j, f)] = 0.0;
        }
        if (VISC) face_state_adjust<NV>(P, X.L, X.R, F, X.fs, var);   // (:265)
    }
}

// cons2prim (ns_time_terms.cc:26-56); with the SA model the row of rho nu~
template <int NV>
__device__ __forceinline__ void cons2prim(const Params& P, const double* q, double Pm[NV][NV]) {
    const double p = q[0], u = q[1], v = q[2], w = q[3], T = q[4];
    double rho = m_rho(P, p, T);
    double pressure = p + P.Pref, temperature = T + P.Tref;
    double drho_dT = -1. * rho / temperature, drho_dp = rho / pressure;
    double c_p = P.Cp;
    double H = c_p * temperature + 0.5 * (u * u + v * v + w * w);
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) Pm[i][j] = 0.0;
    Pm[0][0] = drho_dp; Pm[0][4] = drho_dT;
    Pm[1][0] = drho_dp * u; Pm[1][1] = rho; Pm[1][4] = drho_dT * u;
    Pm[2][0] = drho_dp * v; Pm[2][2] = rho; Pm[2][4] = drho_dT * v;
    Pm[3][0] = drho_dp * w; Pm[3][3] = rho; Pm[3][4] = drho_dT * w;
    Pm[4][0] = drho_dp * H - 1.; Pm[4][1] = rho * u; Pm[4][2] = rho * v; Pm[4][3] = rho * w;
    Pm[4][4] = drho_dT * H + rho * c_p;
    if constexpr (NV == 6) { Pm[5][0] = drho_dp * q[5]; Pm[5][4] = drho_dT * q[5]; Pm[5][5] = rho; }
}

template <int NV>
__global__ void __launch_bounds__(256) diag_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ G,
                                                   const double* __restrict__ UPD, const double* __restrict__ JL,
                                                   const double* __restrict__ JR, const double* __restrict__ dt,
                                                   double* __restrict__ D) {
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld, Fs = g.F;
    constexpr int B2 = NV * NV;
    double blk[B2];
#pragma unroll
    for (int k = 0; k < B2; ++k) blk[k] = 0.0;
    for (int h = g.hf_ptr[c]; h < g.hf_ptr[c + 1]; ++h) {
        int fs = g.hf_face[h];
        if (fs >= 0) {
#pragma unroll
            for (int var = 0; var < NV; ++var)
#pragma unroll
                for (int j = 0; j < NV; ++j) blk[NV * j + var] += -1. * JL[pidx(Fs, NV * var + j, fs)];
        } else {
            int f = ~fs;
#pragma unroll
            for (int var = 0; var < NV; ++var)
#pragma unroll
                for (int j = 0; j < NV; ++j) blk[NV * j + var] += JR[pidx(Fs, NV * var + j, f)];
        }
    }
    double q[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = W[v * ld + c];
    if constexpr (NV == 6) {   // SA source block, gradients frozen
        auto gvec = [&](int v) { return mk<double>(G[pidx(ld, 3 * v, c)], G[pidx(ld, 3 * v + 1, c)], G[pidx(ld, 3 * v + 2, c)]); };
        const Vec<double> gu = gvec(1), gv = gvec(2), gw = gvec(3), gn = gvec(5);
        const double vol = g.vol[c], d = g.wd[c];
        const double S0 = vol * sa_source<double>(m_rho(P, q[0], q[4]), m_visc(P, q[4]), q[5], gu, gv, gw, gn, d);
#pragma unroll
        for (int var = 0; var < NV; ++var) {
            double eps = fd_eps(UPD[var * ld + c], fd_floor(P, var));
            double qp[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) qp[v] = q[v];
            qp[var] += eps;
            double Sp = vol * sa_source<double>(m_rho(P, qp[0], qp[4]), m_visc(P, qp[4]), qp[5], gu, gv, gw, gn, d);
            blk[NV * 5 + var] += -1. * ((Sp - S0) / eps);
        }
    }
    if (dt) {
        double Pm[NV][NV];
        cons2prim<NV>(P, q, Pm);
        double vol = g.vol[c], dtc = dt[c];
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int j = 0; j < NV; ++j) blk[NV * i + j] += Pm[i][j] * vol / dtc;
    }
#pragma unroll
    for (int k = 0; k < B2; ++k) D[pidx(ld, k, c)] = blk[k];
}

template <int FLUX, int VISC, int NV>
static void launch_fdjac_t(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, int has_phi,
                           double phi_const, const double* UPD, double* JL, double* JR, int* status, hipStream_t s) {
    hipLaunchKernelGGL((fdjac_kernel<FLUX, VISC, NV>), dim3((g.F + 127) / 128), dim3(128), 0, s, g, P, W, G, PHI, has_phi,
                       phi_const, UPD, JL, JR, status);
}

template <int FLUX>
static void launch_fdjac_v(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, int has_phi,
                           double phi_const, const double* UPD, double* JL, double* JR, int* status, hipStream_t s) {
    if (g.nv == 6) launch_fdjac_t<FLUX, 1, 6>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s);
    else if (P.kw) launch_fdjac_t<FLUX, kViscRans, 5>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s);
    else if (P.viscous) launch_fdjac_t<FLUX, 1, 5>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s);
    else launch_fdjac_t<FLUX, 0, 5>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s);
}

void launch_fd_jacobian(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI,
                        const double* UPD, const double* dt, double* JL, double* JR, double* D, int* status, hipStream_t s) {
    int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    int has_phi = lim != FCFD_LIMITER_NONE;
    double phi_const = (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0;
    switch (P.flux) {
    case FCFD_FLUX_ROE: launch_fdjac_v<FCFD_FLUX_ROE>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s); break;
    case FCFD_FLUX_VAN_LEER: launch_fdjac_v<FCFD_FLUX_VAN_LEER>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s); break;
    case FCFD_FLUX_SD_SLAU: launch_fdjac_v<FCFD_FLUX_SD_SLAU>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s); break;
    case FCFD_FLUX_SW: launch_fdjac_v<FCFD_FLUX_SW>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s); break;
    default: launch_fdjac_v<FCFD_FLUX_AUSM_PLUS_UP>(g, P, W, G, PHI, has_phi, phi_const, UPD, JL, JR, status, s); break;
    }
    dim3 grid((g.N + 255) / 256), blk(256);
    if (g.nv == 6) hipLaunchKernelGGL(diag_kernel<6>, grid, blk, 0, s, g, P, W, G, UPD, JL, JR, dt, D);
    else hipLaunchKernelGGL(diag_kernel<5>, grid, blk, 0, s, g, P, W, G, UPD, JL, JR, dt, D);
}

}  // namespace fcfd
