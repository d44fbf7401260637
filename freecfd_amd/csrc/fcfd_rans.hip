// Segregated two-equation RANS on the device (src/rans/): k-omega, k-epsilon
// in its omega form, BSL and SST, solved after each NS step as RANS::solve
// does (rans.cc:85-98): face loop and sources (rans_terms.cc:39-333), time
// terms (rans_time_terms.cc:25-86), a GMRES solve of the 2 x 2 block system,
// update_variables with its limits (rans.cc:180-252), then ghosts, eddy
// viscosity (rans_update_eddy_viscosity.cc:25-59) and gradients
// (rans.cc:172-178).
//
// Layout: k and omega as [2][ld] rows (interior cells and ghosts), mu_t and
// the strain rate [ld], gradients of k and omega [6][ld] (k x, y, z, omega
// x, y, z; interior cells).  The face loop writes, per face, its
// contributions to its parent's rows ([6][F]: rhs k, rhs omega, jacL_k,
// jacL_omega, jacR_k, jacR_omega); a cell kernel sums them in reference face
// order, adds the sources and rho V / dt, and writes the right-hand side
// ([2][ld]) and the diagonal 2 x 2 block (row pairs).  The face kernel also
// writes the off-diagonal blocks in the NS operator's convention (row P,
// column N = -JR^T; row N, column P = +JL^T), so the NS Krylov kernels at
// nv = 2 apply the operator.
#include <hip/hip_runtime.h>

#include "fcfd_face.h"
#include "fcfd_rans.h"

namespace fcfd {

namespace {

// model constants (rans.cc:27-41): sigma_k, sigma_omega, beta, beta_star, kappa, alpha
struct KwCoef {
    double sigma_k, sigma_omega, beta, beta_star, alpha;
};
constexpr KwCoef kKomega{0.85, 0.5, 0.075, 0.09, 5. / 9.};
constexpr KwCoef kKeps{1., 0.856, 0.0828, 0.09, 0.44};

// the constants blended with F1 (rans_terms.cc:89-93, 219-223)
__device__ __forceinline__ KwCoef kw_blend(double F1) {
    KwCoef c;
    c.sigma_omega = F1 * kKomega.sigma_omega + (1. - F1) * kKeps.sigma_omega;
    c.sigma_k = F1 * kKomega.sigma_k + (1. - F1) * kKeps.sigma_k;
    c.alpha = F1 * kKomega.alpha + (1. - F1) * kKeps.alpha;
    c.beta = F1 * kKomega.beta + (1. - F1) * kKeps.beta;
    c.beta_star = F1 * kKomega.beta_star + (1. - F1) * kKeps.beta_star;
    return c;
}

// get_blending (rans_terms.cc:316-328): F1
__device__ __forceinline__ double kw_blending(double k, double omega, double rho, double y, double visc, const Vec<double>& gk,
                                              const Vec<double>& go) {
    const double cross_diff = gk.x * go.x + gk.y * go.y + gk.z * go.z;
    const double CD = smax(2. * rho * kKeps.sigma_omega * cross_diff / omega, 1.e-10);
    const double arg1 = 2. * sqrt(k) / (kKeps.beta_star * omega * y);
    const double arg2 = 500. * visc / (rho * y * y * omega);
    const double arg3 = 4. * rho * kKeps.sigma_omega * k / (CD * y * y);
    const double arg = smin(smax(arg1, arg2), arg3);
    return tanh(pow(arg, 4.));
}

__device__ __forceinline__ bool kw_blended(int model) { return model == FCFD_RANS_BSL || model == FCFD_RANS_SST; }
__device__ __forceinline__ double kw_fixed_blend(int model) { return model == FCFD_RANS_KOMEGA ? 1. : 0.; }

// sqrt(2 S_ij S_ij) from the NS velocity gradients of cell c (rans_terms.cc:242-246)
__device__ __forceinline__ double kw_strain(const double* G, size_t ld, int c, double& divU) {
    double gr[12];
    load_pairs<12>(G, ld, c, gr);   // p x y z, u x y z, v x y z, w x y z
    const double dudx = gr[3], dudy = gr[4], dudz = gr[5];
    const double dvdx = gr[6], dvdy = gr[7], dvdz = gr[8];
    const double dwdx = gr[9], dwdy = gr[10], dwdz = gr[11];
    divU = dudx + dvdy + dwdz;
    return sqrt(2. * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + (dudy + dvdx) * (dudy + dvdx) + (dudz + dwdx) * (dudz + dwdx) +
                (dvdz + dwdy) * (dvdz + dwdy));
}

__device__ __forceinline__ double rho_of(const Params& P, const double* W, size_t ld, int c) { return m_rho(P, W[c], W[4 * ld + c]); }

// Variable::face_calculate of a scalar (variable.h:134-147) over the
// two-cell stencil in key order
__device__ __forceinline__ double face_avg(double wp, double wn, double qp, double qn) {
    double t = 0.;
    t += wp * qp;
    t += wn * qn;
    return t;
}

}  // namespace

// ---------------------------------------------------------------------------
// apply_initial_conditions (rans.cc:120-170), one region over the domain:
// k = max(k_low, 1.5 (I |V|)^2), omega = rho k / (ratio mu(T)); gradients 0;
// strain rate 0 on every cell
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kw_init_kernel(DevGrid g, Params P, const double* __restrict__ W, double intensity,
                                                      double visc_ratio, KwDev k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ld = g.ld;
    if (c >= g.N + g.B) return;
    k.strain[c] = 0.;
    if (c >= g.N) return;
    const double u = W[ld + c], v = W[2 * ld + c], w = W[3 * ld + c];
    double kk = intensity * sqrt(u * u + v * v + w * w);
    kk *= 1.5 * kk;
    kk = smax(P.k_low, kk);
    k.K[c] = kk;
    k.K[ld + c] = rho_of(P, W, ld, c) * kk / (visc_ratio * m_visc(P, W[4 * ld + c]));
#pragma unroll
    for (int i = 0; i < 6; ++i) k.GK[i * ld + c] = 0.;
}

// set_bcs (rans_set_bcs.cc:39-73): for each inlet BC, from the NS face
// values of its first face f0[b] (reference order): k = max(k_low, 1.5 (I |V|)^2),
// omega = rho k / (ratio mu(T)), mu_t = ratio mu(T).  out[3 b ..] = k, omega, mu_t.
__global__ void kw_inlet_kernel(DevGrid g, Params P, int n_bc, const double* __restrict__ W, const int* __restrict__ f0,
                                double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_bc) return;
    const BcDev& bc = g.bc[b];
    const int f = f0[b];
    if (bc.type != FCFD_BC_INLET || f < 0) return;
    const size_t ld = g.ld;
    const double2 w = g.fgeo[(size_t)5 * g.F + f];
    const int4 fc = g.fcell[f];
    const int p = fc.x, n = fc.y;
    const double rho = bc.rho_fixed ? bc.rho : face_avg(w.x, w.y, rho_of(P, W, ld, p), rho_of(P, W, ld, n));
    const double T = bc.T_fixed ? bc.T : face_avg(w.x, w.y, W[4 * ld + p], W[4 * ld + n]);
    double V[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) V[i] = bc.V_fixed ? bc.V[i] : face_avg(w.x, w.y, W[(1 + i) * ld + p], W[(1 + i) * ld + n]);
    double kb = bc.kw_turb_intensity * sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    kb *= 1.5 * kb;
    kb = smax(P.k_low, kb);
    const double mu = m_visc(P, T);
    out[3 * b] = kb;
    out[3 * b + 1] = rho * kb / (bc.kw_eddy_visc_ratio * mu);
    out[3 * b + 2] = bc.kw_eddy_visc_ratio * mu;
}

// ghosts: the BC's k and omega where it fixes them (inlets), else the
// parent's (the RANS update_boundaries has no source, rans.h:99; the same
// decision as the NS ghosts)
__global__ void __launch_bounds__(256) kw_ghost_kernel(DevGrid g, double* __restrict__ K) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.B) return;
    const size_t ld = g.ld;
    const int f = g.gh_face[i];
    const int4 fc = g.fcell[f];
    const int gc = g.N + i, p = fc.x;
    const BcDev& bc = g.bc[fc.z];
    K[gc] = bc.kw_kfix ? bc.kw_k : K[p];
    K[ld + gc] = bc.kw_kfix ? bc.kw_om : K[ld + p];
}

// update_eddy_viscosity (rans_update_eddy_viscosity.cc:25-59), interior
// cells and ghosts: SST a1 rho k / max(a1 omega, S F2), else rho k / omega;
// capped at the viscosity ratio limit
__global__ void __launch_bounds__(256) kw_mut_kernel(DevGrid g, Params P, const double* __restrict__ W, KwDev k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N + g.B) return;
    const size_t ld = g.ld;
    const double a1 = 0.31;
    const double mu = m_visc(P, W[4 * ld + c]), rho = rho_of(P, W, ld, c), d = g.wd[c];
    const double kk = k.K[c], om = k.K[ld + c];
    double mt;
    if (P.rans_model == FCFD_RANS_SST) {
        const double arg1 = 2. * sqrt(kk + 1.e-15) / (kKeps.beta_star * om * d);
        const double arg2 = 500. * mu / (rho * om * d * d);
        const double arg3 = smax(arg1, arg2);
        const double F2 = tanh(arg3 * arg3);
        mt = a1 * rho * kk / smax(a1 * om, k.strain[c] * F2);
    } else {
        mt = rho * kk / om;
    }
    k.mut[c] = smin(P.visc_ratio_limit * mu, mt);
}

// calc_cell_grads (rans.cc:172-178) with Variable::cell_gradient
// (variable.h:193-216): the gradient map (sliced ELL, entries in key order)
// or, for a cell without one, the Green-Gauss face sum of the face values
// (the BC's value where k and omega are fixed, else the face average)
__global__ void __launch_bounds__(256) kw_grad_kernel(DevGrid g, KwDev k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    double a[3] = {0., 0., 0.}, b[3] = {0., 0., 0.};
    if (g.ggc && g.ggc[c]) {
        const double vol = g.vol[c];
        for (int h = g.hf_ptr[c]; h < g.hf_ptr[c + 1]; ++h) {
            const int fs = g.hf_face[h];
            const int f = fs >= 0 ? fs : ~fs;
            const double A = g.farea[f];
            double a0 = g.fnx[f] * A / vol, a1 = g.fny[f] * A / vol, a2 = g.fnz[f] * A / vol;
            if (fs < 0) { a0 = a0 * -1.; a1 = a1 * -1.; a2 = a2 * -1.; }
            const int4 fc = g.fcell[f];
            double fk, fo;
            if (fc.z >= 0 && g.bc[fc.z].kw_kfix) {
                fk = g.bc[fc.z].kw_k;
                fo = g.bc[fc.z].kw_om;
            } else {
                const double2 w = g.ggw[f];
                fk = face_avg(w.x, w.y, k.K[fc.x], k.K[fc.y]);
                fo = face_avg(w.x, w.y, k.K[ld + fc.x], k.K[ld + fc.y]);
            }
            a[0] += fk * a0; a[1] += fk * a1; a[2] += fk * a2;
            b[0] += fo * a0; b[1] += fo * a1; b[2] += fo * a2;
        }
    } else {
        const int s = c >> 6, lane = c & 63;
        const int e0 = g.gslice[s], K = (g.gslice[s + 1] - e0) >> 6;
        for (int e = 0; e < K; ++e) {
            const double2 wa = g.gwa[e0 + 64 * e + lane], wb = g.gwb[e0 + 64 * e + lane];
            const int j = (int)__double_as_longlong(wb.y);
            const double qk = k.K[j], qo = k.K[ld + j];
            a[0] += wa.x * qk; a[1] += wa.y * qk; a[2] += wb.x * qk;
            b[0] += wa.x * qo; b[1] += wa.y * qo; b[2] += wb.x * qo;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        k.GK[i * ld + c] = a[i];
        k.GK[(3 + i) * ld + c] = b[i];
    }
}

// ---------------------------------------------------------------------------
// RANS::terms face loop (rans_terms.cc:51-187, get_kOmega :271-314): mdot and
// weightL of the NS face loop at the same state (the NS face evaluation is
// recomputed here), convective and diffusive fluxes of k and omega, and the
// implicit coefficients.  FV [6][F]: the parent's rhs contributions
// (diffusive - convective) and jacL_k, jacL_omega, jacR_k, jacR_omega.  JL,
// JR (2 x 2 blocks, row pairs): -diag(jacL), -diag(jacR), so that the NS
// operator convention gives row P, column N = jacR and row N, column P =
// -jacL.  lam_last: the laminar viscosity of the reference's last face
// (the cell loop reads the face loop's last lam_visc, rans_terms.cc:183-185).
// ---------------------------------------------------------------------------
template <int FLUX>
__global__ void __launch_bounds__(64) kw_face_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ G,
                                                     const double* __restrict__ PHI, int has_phi, double phi_const, KwDev k,
                                                     double* __restrict__ FV, double* __restrict__ JL, double* __restrict__ JR,
                                                     int last_face, double* __restrict__ lam_last, int* __restrict__ status) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.F) return;
    const size_t ld = g.ld, Fs = g.F;
    FaceCtx<double, 5> X;
    CellLoader<5> cl{W, G, PHI, has_phi != 0, phi_const, ld};
    face_base<FLUX, kViscRans, 5>(g, P, cl, f, X, status);
    const FaceGeo& F = X.F;
    const int parent = X.p, neighbor = X.n, b = F.bc;
    const double area = F.area;
    const double mdot = (X.C[0] - X.D[0]) / area;   // mdot.face (ns_assemble_linear_system.cc:122)
    // weightL.face as the flux function leaves it
    const double weightL = weight_left<FLUX>(X.L, X.R, X.C[0]);
    const double weightR = 1. - weightL;
    const double2 wgt = g.fgeo[(size_t)5 * Fs + f];
    const double lam_visc = m_visc(P, (b >= 0 && g.bc[b].T_fixed) ? g.bc[b].T : face_avg(wgt.x, wgt.y, W[4 * ld + parent], W[4 * ld + neighbor]));
    const double turb_visc = X.mut_face;
    const double leftK = k.K[parent], leftOmega = k.K[ld + parent], rightK = k.K[neighbor], rightOmega = k.K[ld + neighbor];
    bool extrapolated = false;
    Vec<double> gK, gO;
    const Vec<double> gKp = mk<double>(k.GK[parent], k.GK[ld + parent], k.GK[2 * ld + parent]);
    const Vec<double> gOp = mk<double>(k.GK[3 * ld + parent], k.GK[4 * ld + parent], k.GK[5 * ld + parent]);
    double faceK, faceOmega, faceRho;
    if (b >= 0) {
        const BcDev& bc = g.bc[b];
        const int ty = bc.type;
        if (ty == FCFD_BC_SYMMETRY || ty == FCFD_BC_OUTLET) extrapolated = true;
        else if (ty == FCFD_BC_WALL && bc.kind == FCFD_KIND_SLIP) extrapolated = true;
        gK = gKp;
        gO = gOp;
        faceK = bc.kw_kfix ? bc.kw_k : face_avg(wgt.x, wgt.y, leftK, rightK);
        faceOmega = bc.kw_kfix ? bc.kw_om : face_avg(wgt.x, wgt.y, leftOmega, rightOmega);
        faceRho = bc.rho_fixed ? bc.rho : face_avg(wgt.x, wgt.y, rho_of(P, W, ld, parent), rho_of(P, W, ld, neighbor));
    } else {
        const Vec<double> gKn = mk<double>(k.GK[neighbor], k.GK[ld + neighbor], k.GK[2 * ld + neighbor]);
        const Vec<double> gOn = mk<double>(k.GK[3 * ld + neighbor], k.GK[4 * ld + neighbor], k.GK[5 * ld + neighbor]);
        gK = vadd(vadd(mk<double>(0., 0., 0.), vsc(wgt.x, gKp)), vsc(wgt.y, gKn));
        gO = vadd(vadd(mk<double>(0., 0., 0.), vsc(wgt.x, gOp)), vsc(wgt.y, gOn));
        faceK = face_avg(wgt.x, wgt.y, leftK, rightK);
        faceOmega = face_avg(wgt.x, wgt.y, leftOmega, rightOmega);
        faceRho = face_avg(wgt.x, wgt.y, rho_of(P, W, ld, parent), rho_of(P, W, ld, neighbor));
    }
    const Vec<double> l2rn = vnorm(F.l2r);
    const double l2rmag = vmag(F.l2r);
    gK = correct(gK, l2rn, l2rmag, rightK - leftK);
    gO = correct(gO, l2rn, l2rmag, rightOmega - leftOmega);
    const double conv0 = mdot * (weightL * leftK + weightR * rightK) * area;
    const double conv1 = mdot * (weightL * leftOmega + weightR * rightOmega) * area;
    double blending = kw_fixed_blend(P.rans_model);
    if (kw_blended(P.rans_model)) blending = kw_blending(faceK, faceOmega, faceRho, g.wd[parent], lam_visc, gK, gO);
    const KwCoef m = kw_blend(blending);
    const double dn = vdot(gK, F.n), dno = vdot(gO, F.n);
    const double diff0 = (lam_visc + turb_visc * m.sigma_k) * dn * area;
    const double diff1 = (lam_visc + turb_visc * m.sigma_omega) * dno * area;
    const double AoverH = area * vdot(F.n, l2rn) / l2rmag;
    double jl0 = weightL * mdot * area, jl1 = weightL * mdot * area;
    if (extrapolated) { jl0 += weightR * mdot * area; jl1 += weightR * mdot * area; }
    if (!extrapolated) { jl0 += (lam_visc + turb_visc * m.sigma_k) * AoverH; jl1 += (lam_visc + turb_visc * m.sigma_omega) * AoverH; }
    double jr0 = 0., jr1 = 0.;
    if (b < 0) {
        jr0 = weightR * mdot * area;
        jr0 -= (lam_visc + turb_visc * m.sigma_k) * AoverH;
        jr1 = weightR * mdot * area;
        jr1 -= (lam_visc + turb_visc * m.sigma_omega) * AoverH;
    }
    FV[f] = diff0 - conv0;
    FV[Fs + f] = diff1 - conv1;
    FV[2 * Fs + f] = jl0;
    FV[3 * Fs + f] = jl1;
    FV[4 * Fs + f] = jr0;
    FV[5 * Fs + f] = jr1;
    double2* JL2 = reinterpret_cast<double2*>(JL);
    double2* JR2 = reinterpret_cast<double2*>(JR);
    JL2[f] = make_double2(-jl0, 0.);
    JL2[Fs + f] = make_double2(0., -jl1);
    JR2[f] = make_double2(-jr0, 0.);
    JR2[Fs + f] = make_double2(0., -jr1);
    if (f == last_face) *lam_last = lam_visc;
}

// the strain rate of every ghost (rans_terms.cc:242-249 run over all cells;
// a ghost's NS gradients are its parent's copies)
__global__ void __launch_bounds__(256) kw_ghost_strain_kernel(DevGrid g, const double* __restrict__ G, double* __restrict__ strain) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.B) return;
    double divU;
    strain[g.N + i] = kw_strain(G, g.ld, g.N + i, divU);
}

// the cell loop of RANS::terms (rans_terms.cc:189-292) and the time terms
// (rans_time_terms.cc:65-66): per interior cell, the faces' contributions in
// ascending reference face order (the face loop's ADD_VALUES order), then
// the sources, then rho V / dt on the diagonal.  Sets the strain rate.
__global__ void __launch_bounds__(256) kw_cell_kernel(DevGrid g, Params P, const double* __restrict__ W, const double* __restrict__ G,
                                                      const double* __restrict__ dt, const double* __restrict__ FV, KwDev k,
                                                      const double* __restrict__ lam_last, double* __restrict__ rhs, double* __restrict__ D) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld, Fs = g.F;
    double divU;
    const double sr = kw_strain(G, ld, c, divU);
    k.strain[c] = sr;
    double r0 = 0., r1 = 0., d00 = 0., d11 = 0.;
    for (int h = g.hf_ptr[c]; h < g.hf_ptr[c + 1]; ++h) {
        const int fs = g.hf_face[h];
        if (fs >= 0) {
            r0 += FV[fs]; r1 += FV[Fs + fs]; d00 += FV[2 * Fs + fs]; d11 += FV[3 * Fs + fs];
        } else {
            const int f = ~fs;
            r0 += -1. * FV[f]; r1 += -1. * FV[Fs + f]; d00 += -FV[4 * Fs + f]; d11 += -FV[5 * Fs + f];
        }
    }
    const double rho = rho_of(P, W, ld, c);
    const double kk = k.K[c], om = k.K[ld + c], mu_t = k.mut[c], vol = g.vol[c];
    const Vec<double> gk = mk<double>(k.GK[c], k.GK[ld + c], k.GK[2 * ld + c]);
    const Vec<double> go = mk<double>(k.GK[3 * ld + c], k.GK[4 * ld + c], k.GK[5 * ld + c]);
    double F1 = kw_fixed_blend(P.rans_model);
    if (kw_blended(P.rans_model)) F1 = kw_blending(kk, om, rho, g.wd[c], *lam_last, gk, go);
    const KwCoef m = kw_blend(F1);
    double Prod_k = mu_t * (sr * sr - 2. / 3. * divU * divU) - 2. / 3. * rho * kk * divU;
    const double Dest_k = m.beta_star * rho * om * kk;
    Prod_k = smin(Prod_k, 10. * Dest_k);
    double cross_diffusion = 0.;
    if (P.rans_model != FCFD_RANS_KOMEGA) cross_diffusion = gk.x * go.x + gk.y * go.y + gk.z * go.z;
    double s0 = Prod_k;
    s0 -= Dest_k;
    double s1 = m.alpha * rho / mu_t * Prod_k;
    s1 -= m.beta * rho * om * om;
    s1 += 2. * (1. - F1) * rho * kKomega.sigma_omega * cross_diffusion / om;
    const double j00 = m.beta_star * rho * om * vol, j01 = m.beta_star * rho * kk * vol;
    double j11 = 2. * m.beta * rho * om * vol;
    j11 += 2. * (1. - F1) * kKomega.sigma_omega * kKomega.sigma_omega * cross_diffusion / (om * om) * vol;
    r0 += s0 * vol;
    r1 += s1 * vol;
    d00 += j00;
    d11 += j11;
    const double tt = rho * vol / dt[c];
    d00 += tt;
    d11 += tt;
    rhs[c] = r0;
    rhs[ld + c] = r1;
    double2* D2 = reinterpret_cast<double2*>(D);
    D2[c] = make_double2(d00, j01);
    D2[ld + c] = make_double2(0., d11);
}

// update_variables (rans.cc:180-252), per interior cell once the NaN/Inf
// count is known to be 0: the limits, k += dk, omega += domega
__global__ void __launch_bounds__(256) kw_update_kernel(int N, size_t ld, Params P, double* __restrict__ K, const double* __restrict__ dU) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    double u0 = dU[c], u1 = dU[ld + c];
    const double kk = K[c], om = K[ld + c];
    u0 = smax(-1. * (kk - P.k_low), u0);
    u0 = smin((P.k_high - kk), u0);
    u1 = smax(-1. * (om - P.omega_low), u1);
    K[c] = kk + u0;
    K[ld + c] = om + u1;
}

// the update's norms (rans.cc:206-222) with the limits applied, and a
// NaN/Inf count of the raw update: partial sums per block, fixed-order tree
constexpr int kKwRedBlocks = 256, kKwRedThreads = 256;
__global__ void __launch_bounds__(kKwRedThreads) kw_norms_kernel(int N, size_t ld, Params P, const double* __restrict__ K,
                                                                 const double* __restrict__ dU, const double* __restrict__ dt,
                                                                 double* __restrict__ part) {
    __shared__ double sh[3][kKwRedThreads];
    double acc[3] = {0., 0., 0.};
    for (int c = blockIdx.x * kKwRedThreads + threadIdx.x; c < N; c += kKwRedBlocks * kKwRedThreads) {
        double u0 = dU[c], u1 = dU[ld + c];
        if (!isfinite(u0) || !isfinite(u1)) { acc[2] += 1.; continue; }
        const double kk = K[c], om = K[ld + c];
        u0 = smax(-1. * (kk - P.k_low), u0);
        u0 = smin((P.k_high - kk), u0);
        u1 = smax(-1. * (om - P.omega_low), u1);
        const double dt2 = dt[c] * dt[c];
        acc[0] += u0 * u0 / dt2;
        acc[1] += u1 * u1 / dt2;
    }
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = kKwRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[q * kKwRedBlocks + blockIdx.x] = sh[q][0];
}
__global__ void __launch_bounds__(kKwRedThreads) kw_norms_final_kernel(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[kKwRedThreads];
    const int q = blockIdx.x;
    sh[threadIdx.x] = part[q * kKwRedBlocks + threadIdx.x];
    __syncthreads();
    for (int s = kKwRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static dim3 blocks(int n, int t) { return dim3((unsigned)((n + t - 1) / t)); }

void kw_launch_init(const DevGrid& g, const Params& P, const double* W, double intensity, double visc_ratio, const KwDev& k,
                    hipStream_t s) {
    hipLaunchKernelGGL(kw_init_kernel, blocks(g.N + g.B, 256), dim3(256), 0, s, g, P, W, intensity, visc_ratio, k);
}

void kw_launch_inlets(const DevGrid& g, const Params& P, int n_bc, const double* W, const int* f0, double* out, hipStream_t s) {
    if (n_bc > 0) hipLaunchKernelGGL(kw_inlet_kernel, blocks(n_bc, 64), dim3(64), 0, s, g, P, n_bc, W, f0, out);
}

void kw_launch_refresh(const DevGrid& g, const Params& P, const double* W, const KwDev& k, hipStream_t s) {
    if (g.B > 0) hipLaunchKernelGGL(kw_ghost_kernel, blocks(g.B, 256), dim3(256), 0, s, g, k.K);
    hipLaunchKernelGGL(kw_mut_kernel, blocks(g.N + g.B, 256), dim3(256), 0, s, g, P, W, k);
    hipLaunchKernelGGL(kw_grad_kernel, blocks(g.N, 256), dim3(256), 0, s, g, k);
}

template <int FLUX>
static void face_launch(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, int has_phi,
                        double phi_const, const KwDev& k, double* FV, double* JL, double* JR, int last_face, double* lam_last,
                        int* status, hipStream_t s) {
    hipLaunchKernelGGL((kw_face_kernel<FLUX>), blocks(g.F, 64), dim3(64), 0, s, g, P, W, G, PHI, has_phi, phi_const, k, FV, JL, JR,
                       last_face, lam_last, status);
}

void kw_launch_terms(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, const double* dt,
                     const KwDev& k, const KwTerms& t, int last_face, int* status, hipStream_t s) {
    const int lim = (P.order == FCFD_ORDER_FIRST) ? FCFD_LIMITER_NONE : P.limiter;
    const int has_phi = lim != FCFD_LIMITER_NONE;
    const double phi_const = (P.order == FCFD_ORDER_FIRST) ? 0.0 : 1.0;
    switch (P.flux) {
    case FCFD_FLUX_ROE: face_launch<FCFD_FLUX_ROE>(g, P, W, G, PHI, has_phi, phi_const, k, t.FV, t.JL, t.JR, last_face, t.lam_last, status, s); break;
    case FCFD_FLUX_VAN_LEER: face_launch<FCFD_FLUX_VAN_LEER>(g, P, W, G, PHI, has_phi, phi_const, k, t.FV, t.JL, t.JR, last_face, t.lam_last, status, s); break;
    case FCFD_FLUX_SD_SLAU: face_launch<FCFD_FLUX_SD_SLAU>(g, P, W, G, PHI, has_phi, phi_const, k, t.FV, t.JL, t.JR, last_face, t.lam_last, status, s); break;
    case FCFD_FLUX_SW: face_launch<FCFD_FLUX_SW>(g, P, W, G, PHI, has_phi, phi_const, k, t.FV, t.JL, t.JR, last_face, t.lam_last, status, s); break;
    default: face_launch<FCFD_FLUX_AUSM_PLUS_UP>(g, P, W, G, PHI, has_phi, phi_const, k, t.FV, t.JL, t.JR, last_face, t.lam_last, status, s); break;
    }
    if (g.B > 0) hipLaunchKernelGGL(kw_ghost_strain_kernel, blocks(g.B, 256), dim3(256), 0, s, g, G, k.strain);
    hipLaunchKernelGGL(kw_cell_kernel, blocks(g.N, 256), dim3(256), 0, s, g, P, W, G, dt, t.FV, k, t.lam_last, t.rhs, t.D);
}

hipError_t kw_update(const DevGrid& g, const Params& P, const KwDev& k, const double* dU, const double* dt, double* part,
                     double* out_dev, double out_host[3], hipStream_t s) {
    hipLaunchKernelGGL(kw_norms_kernel, dim3(kKwRedBlocks), dim3(kKwRedThreads), 0, s, g.N, (size_t)g.ld, P, k.K, dU, dt, part);
    hipLaunchKernelGGL(kw_norms_final_kernel, dim3(3), dim3(kKwRedThreads), 0, s, part, out_dev);
    hipError_t e = hipMemcpyAsync(out_host, out_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (out_host[2] != 0.0) return hipSuccess;   // NaN/Inf: nothing changes
    hipLaunchKernelGGL(kw_update_kernel, blocks(g.N, 256), dim3(256), 0, s, g.N, (size_t)g.ld, P, k.K, dU);
    return hipGetLastError();
}

}  // namespace fcfd
