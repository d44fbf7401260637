// C-ABI entry points of libfcfd (declared in include/fcfd_gpu.h): context
// creation (host set-up + device upload), layout helpers and the residual.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fcfd_adjoint.h"
#include "fcfd_device.h"
#include "fcfd_internal.h"
#include "fcfd_jacobian.h"
#include "fcfd_krylov.h"
#include "fcfd_mg.h"
#include "fcfd_rans.h"
#include "fcfd_tangent.h"

#include <chrono>
#include <mutex>

using namespace fcfd;

struct fcfd_ctx {
    HostGrid hg;
    fcfd_options opt;
    bool gpu_wall_distance = false;   // wall distances left to the device set-up (large meshes)
    Params P;
    std::vector<BcDev> bch;
    bool device = false;
    DevGrid dg{};
    std::vector<void*> allocs;
    double *G = nullptr, *PHI = nullptr, *dref = nullptr, *FX = nullptr;
    double* SRC = nullptr;   // [ld] V S of the SA source (nv = 6), from grad_lim
    double* adj_v = nullptr; // [nv][ld] J^T x inside the adjoint step's operator
    Multigrid mg;            // aggregation multigrid hierarchy (built on first use)
    bool mg_tried = false;
    // the V-cycle as a HIP graph (fine_in -> fine_out), captured once on a
    // private stream and replayed on the caller's; FCFD_NO_GRAPHS=1 launches
    // the same kernels one by one instead
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t mg_exec = nullptr;
    int mg_exec_sweeps = -1;
    int* status = nullptr;
    double *JL = nullptr, *JR = nullptr, *Dg = nullptr, *zero_upd = nullptr;   // FD Jacobian (lazy)
    // Newton-Krylov workspace (lazy)
    double *R = nullptr, *dt = nullptr, *dW = nullptr, *upd = nullptr, *red_part = nullptr, *red_out = nullptr;
    double* dt_red = nullptr;   // 2 + 256 doubles: CFL_MAX step, largest CFL, block partials (lazy)
    // dual time stepping (lazy): pseudo step size, drho/dp of the pseudo
    // block, the interior cells' neighbour cells (WS95)
    double *dtau = nullptr, *drdp = nullptr;
    const int *nb_ptr = nullptr, *nb_idx = nullptr;
    int gm_restart32 = 0, gm_restart64 = 0;   // restart length of the allocated FP32 / FP64 Krylov bases
    GmresWork<double> w64;
    GmresWork<float> w32;
    double* Dinv64 = nullptr;
    float *Dinv32 = nullptr, *D32 = nullptr, *JL32 = nullptr, *JR32 = nullptr, *b32 = nullptr, *x32 = nullptr;
    // tangent workspace (lazy)
    double *dG = nullptr, *dPHI = nullptr, *dFX = nullptr, *jv_in = nullptr, *jv_out = nullptr, *jv_R = nullptr, *dSRC = nullptr;
    // the FD Jacobian's off-diagonal blocks re-laid for the streamed matvec (lazy)
    SellBlocks sell;
    double* sell64 = nullptr;
    float* sell32 = nullptr;
    // FP32 tangent workspace of the FP32 matrix-free Newton operator (lazy)
    float *dG32 = nullptr, *dPHI32 = nullptr, *dFX32 = nullptr, *jv_in32 = nullptr, *jv_R32 = nullptr, *dSRC32 = nullptr;
    LimLin lin32{nullptr, nullptr, nullptr};   // the limiters' linearisation for the FP32 operator (limited cases)
    // cell colouring for the block Gauss-Seidel preconditioner (lazy)
    Colors colors;
    // adjoint and loads (lazy)
    bool adj_ready = false;
    AdjGrid ag{};
    AdjWork aw{};
    int n_loadfaces = 0;
    double *loads_tmp = nullptr, *loads_out = nullptr;
    // segregated two-equation RANS (opt.turbulence = FCFD_TURB_KW): fields,
    // terms, Krylov workspace at nv = 2, the inlets' k, omega, mu_t
    bool kw_ready = false;            // fcfd_rans_init or fcfd_rans_set has run
    KwDev kw{};
    KwTerms kwt{};
    double *kw_dU = nullptr, *kw_part = nullptr, *kw_out = nullptr, *kw_inl = nullptr, *kw_Dinv = nullptr;
    const int* kw_f0 = nullptr;       // per BC: internal id of its first face in reference order (-1: none)
    std::vector<double> kw_bcv;       // 3 per BC: k, omega, mu_t (inlets)
    GmresWork<double> kw_w;
    int kw_restart = 0;
    fcfd_rans_opts kw_opts{30, 10, 1e-6, 1e-12};   // used inside fcfd_newton_step (the reference's defaults)
    std::string err;
    // optional per-stage timing: a ring of event quadruples recorded on the
    // launch stream around the three kernels, folded in lazily so that
    // timing adds no host synchronisation to the call sequence
    static const int kRing = 512;
    bool profiling = false;
    static const int kStages = 4;     // ghosts, gradients+limiters, face fluxes, cell sums
    std::vector<hipEvent_t> ev;       // (kStages+1) * kRing
    int ring_head = 0, ring_count = 0;
    double stage_ms[kStages] = {0.0, 0.0, 0.0, 0.0};
    long long stage_calls = 0;
};

namespace {

int fail(fcfd_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(ctx, FCFD_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int upload(fcfd_ctx* c, const std::vector<T>& h, const T** out, size_t pad_to = 0) {
    size_t n = std::max(h.size(), pad_to);
    if (n == 0) n = 1;
    void* p = nullptr;
    HIPCHK(c, hipMalloc(&p, n * sizeof(T)));
    c->allocs.push_back(p);
    if (n > h.size()) HIPCHK(c, hipMemset(p, 0, n * sizeof(T)));
    if (!h.empty()) HIPCHK(c, hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T*)p;
    return FCFD_OK;
}

// release one allocation of the context (nullptr: nothing)
void dfree(fcfd_ctx* c, void* p) {
    if (!p) return;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
    if (it != c->allocs.end()) {
        (void)hipFree(p);
        c->allocs.erase(it);
    }
}

template <typename T>
int dalloc(fcfd_ctx* c, size_t n, T** out) {
    void* p = nullptr;
    HIPCHK(c, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    c->allocs.push_back(p);
    HIPCHK(c, hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    *out = (T*)p;
    return FCFD_OK;
}

int validate_options(fcfd_ctx* c) {
    const fcfd_options& o = c->opt;
    if (o.gamma <= 1.0 || o.Mw <= 0.0) return fail(c, FCFD_EINVAL, "material: gamma must exceed 1 and Mw be positive");
    for (int b = 0; b < o.n_bc; ++b) {
        const fcfd_bc& bc = o.bc[b];
        if (bc.type < FCFD_BC_SYMMETRY || bc.type > FCFD_BC_OUTLET) return fail(c, FCFD_EBC, "BC_" + std::to_string(b + 1) + ": unknown type");
        if (bc.type == FCFD_BC_OUTLET && bc.specified == FCFD_SPEC_STATE)   // ns_apply_bcs.cc:165-167
            return fail(c, FCFD_EBC, "You can't specify complete thermodynamic state at outlet BC_" + std::to_string(b + 1));
        if (bc.type == FCFD_BC_INLET && (bc.kind == FCFD_KIND_VELOCITY || bc.kind == FCFD_KIND_MDOT) &&
            bc.specified == FCFD_SPEC_NONE)   // ns_apply_bcs.cc:78-81,123-126
            return fail(c, FCFD_EBC, "At least one of p,T,rho needs to be specified at Inlet at BC_" + std::to_string(b + 1));
    }
    if (o.turbulence != FCFD_TURB_NONE && o.turbulence != FCFD_TURB_SA && o.turbulence != FCFD_TURB_KW)
        return fail(c, FCFD_EINVAL, "unknown turbulence model");
    if (o.turbulence == FCFD_TURB_KW) {
        if (o.rans_model < FCFD_RANS_KEPSILON || o.rans_model > FCFD_RANS_SST) return fail(c, FCFD_EINVAL, "unknown two-equation RANS model");
        if (!(o.visc_model == FCFD_VISC_SUTHERLAND || o.mu > 0.0))
            return fail(c, FCFD_EINVAL, "the two-equation RANS model needs a positive molecular viscosity");
        if (!(o.Pr_t > 0.0)) return fail(c, FCFD_EINVAL, "the two-equation RANS model needs a positive turbulent Prandtl number");
        if (!(o.omega_low > 0.0) || !(o.k_high > o.k_low)) return fail(c, FCFD_EINVAL, "two-equation RANS limits: omega_low > 0, k_high > k_low");
    }
    if (o.turbulence == FCFD_TURB_SA) {
        if (!(o.visc_model == FCFD_VISC_SUTHERLAND || o.mu > 0.0))
            return fail(c, FCFD_EINVAL, "the Spalart-Allmaras model needs a positive molecular viscosity");
        if (!(o.Pr_t > 0.0)) return fail(c, FCFD_EINVAL, "the Spalart-Allmaras model needs a positive turbulent Prandtl number");
    }
    return FCFD_OK;
}

void set_params(fcfd_ctx* c) {
    const fcfd_options& o = c->opt;
    Params& P = c->P;
    P.gamma = o.gamma;
    P.R = 8314.47215 / o.Mw;
    P.Cp = o.gamma * P.R / (o.gamma - 1.);
    P.Pref = o.Pref; P.Tref = o.Tref;
    P.visc_model = o.visc_model; P.mu = o.mu; P.sut_mu_ref = o.sut_mu_ref; P.sut_T_ref = o.sut_T_ref; P.sut_S = o.sut_S;
    P.cond_model = o.cond_model; P.lambda = o.lambda; P.Pr = o.Pr;
    P.order = o.order; P.limiter = o.limiter; P.flux = o.convective_flux;
    P.limiter_threshold = o.limiter_threshold; P.Minf = o.Minf; P.wdiss = o.wdiss; P.bl_height = o.bl_height;
    P.glscale = c->hg.glscale;
    bool inviscid = (o.visc_model == FCFD_VISC_CONSTANT && o.mu == 0.0) &&
                    (o.cond_model == FCFD_COND_CONSTANT && o.lambda == 0.0);
    bool fixedq = false;
    for (int b = 0; b < o.n_bc; ++b) fixedq |= (o.bc[b].thermal == FCFD_THERMAL_FIXED_Q && o.bc[b].qdot != 0.0);
    P.viscous = (inviscid && !fixedq) ? 0 : 1;
    P.nv = (o.turbulence == FCFD_TURB_SA) ? 6 : 5;
    P.Pr_t = o.Pr_t;
    P.kw = o.turbulence == FCFD_TURB_KW;
    P.seed = BcSeed{};
    P.seed.index = -1;
    P.rans_model = o.rans_model;
    P.k_low = o.k_low; P.k_high = o.k_high; P.omega_low = o.omega_low; P.visc_ratio_limit = o.visc_ratio_limit;
    double nut_scale = 0.0;
    for (int b = 0; b < o.n_bc; ++b) nut_scale = std::max(nut_scale, std::fabs(o.bc[b].nut));
    P.nut_eps = 10. * 1.4901161193847656e-08 * (nut_scale > 0.0 ? nut_scale : 1.0);
    c->bch.assign(std::max(o.n_bc, 1), BcDev{});
    for (int b = 0; b < o.n_bc; ++b) {
        const fcfd_bc& s = o.bc[b];
        BcDev& d = c->bch[b];
        d.type = s.type; d.kind = s.kind; d.specified = s.specified; d.thermal = s.thermal;
        d.p = s.p; d.T = s.T; d.rho = s.rho; d.V[0] = s.V[0]; d.V[1] = s.V[1]; d.V[2] = s.V[2];
        d.p_total = s.p_total; d.T_total = s.T_total; d.mdot = s.mdot; d.qdot = s.qdot;
        d.total_area = b < (int)c->hg.bc_area.size() ? c->hg.bc_area[b] : 0.0;
        d.nut = s.nut;
        d.p_fixed = s.p_fixed; d.T_fixed = s.T_fixed; d.V_fixed = s.V_fixed; d.qdot_fixed = s.qdot_fixed;
        d.rho_fixed = s.rho_fixed;
        // two-equation RANS (rans_set_bcs.cc:39-79): inlets fix k, omega and
        // mu_t (values from fcfd_rans_init), no-slip walls fix mu_t = 0
        d.kw_kfix = s.type == FCFD_BC_INLET;
        d.kw_mfix = s.type == FCFD_BC_INLET || (s.type == FCFD_BC_WALL && s.kind != FCFD_KIND_SLIP);
        const bool have = (int)c->kw_bcv.size() >= 3 * (b + 1);
        d.kw_k = (d.kw_kfix && have) ? c->kw_bcv[3 * b] : 0.0;
        d.kw_om = (d.kw_kfix && have) ? c->kw_bcv[3 * b + 1] : 0.0;
        d.kw_mut = (s.type == FCFD_BC_INLET && have) ? c->kw_bcv[3 * b + 2] : 0.0;
        d.kw_turb_intensity = s.turb_intensity; d.kw_eddy_visc_ratio = s.eddy_visc_ratio;
    }
}

// Half-face lists in internal numbering: the faces of internal cell i are
// hf[hp[i] .. hp[i+1]), in ascending reference face order, each the internal
// face id when i is the face's parent and its complement ~id otherwise.
void half_faces(const HostGrid& g, std::vector<int>& hp, std::vector<int>& hf) {
    const int N = g.N;
    hp.assign(N + 1, 0);
    hf.clear();
    hf.reserve(g.cfaces.size());
    for (int i = 0; i < N; ++i) {
        int r = g.iperm[i];
        for (int k = g.cfptr[r]; k < g.cfptr[r + 1]; ++k) {
            int f = g.cfaces[k];
            int fi = g.fperm[f];
            hf.push_back(g.fparent[f] == r ? fi : ~fi);
        }
        hp[i + 1] = (int)hf.size();
    }
}

int build_device(fcfd_ctx* c) {
    HostGrid& g = c->hg;
    const int N = g.N, B = g.B, F = g.F, T = N + B;
    DevGrid& d = c->dg;
    d.N = N; d.B = B; d.F = F;
    d.nv = c->P.nv;
    d.ld = ((T + 63) / 64) * 64;
    const int ld = d.ld;
    d.ref_first = g.perm[0];
    d.ref_last = g.perm[N - 1];
    // cells
    {
        std::vector<double> x(ld, 0.0), y(ld, 0.0), z(ld, 0.0), v(ld, 0.0), l(ld, 0.0), w(ld, 1.e20);
        for (int i = 0; i < T; ++i) {
            int r = g.iperm[i];
            x[i] = g.cc[3 * r]; y[i] = g.cc[3 * r + 1]; z[i] = g.cc[3 * r + 2];
            v[i] = g.vol[r]; l[i] = g.lscale[r]; w[i] = g.wdist[r];
        }
        int rc;
        if ((rc = upload(c, x, &d.ccx)) || (rc = upload(c, y, &d.ccy)) || (rc = upload(c, z, &d.ccz)) ||
            (rc = upload(c, v, &d.vol)) || (rc = upload(c, l, &d.lscale)) || (rc = upload(c, w, &d.wd)))
            return rc;
        std::vector<double2> c2((size_t)2 * ld);
        for (int i = 0; i < ld; ++i) {
            c2[i] = make_double2(x[i], y[i]);
            c2[(size_t)ld + i] = make_double2(z[i], v[i]);
        }
        if ((rc = upload(c, c2, &d.cc2))) return rc;
    }
    // gradient maps, sliced ELL with 64-cell slices
    {
        int ns = (N + 63) / 64;
        std::vector<int> gs(ns + 1, 0), gi;
        std::vector<double> wx, wy, wz;
        for (int s = 0; s < ns; ++s) {
            int width = 0;
            for (int lane = 0; lane < 64; ++lane) {
                int i = s * 64 + lane;
                if (i < N) { int r = g.iperm[i]; width = std::max(width, g.gptr[r + 1] - g.gptr[r]); }
            }
            size_t base = gi.size();
            gi.resize(base + (size_t)width * 64);
            wx.resize(gi.size(), 0.0); wy.resize(gi.size(), 0.0); wz.resize(gi.size(), 0.0);
            for (int lane = 0; lane < 64; ++lane) {
                int i = s * 64 + lane;
                int r = i < N ? g.iperm[i] : -1;
                int cnt = r >= 0 ? g.gptr[r + 1] - g.gptr[r] : 0;
                for (int k = 0; k < width; ++k) {
                    size_t e = base + (size_t)k * 64 + lane;
                    if (k < cnt) {
                        int ent = g.gptr[r] + k;
                        gi[e] = g.perm[g.gidx[ent]];
                        wx[e] = g.gw[3 * ent]; wy[e] = g.gw[3 * ent + 1]; wz[e] = g.gw[3 * ent + 2];
                    } else {
                        gi[e] = i < N ? i : 0;   // zero-weight pad
                    }
                }
            }
            gs[s + 1] = (int)gi.size();
        }
        std::vector<double2> ga(gi.size()), gb(gi.size());
        for (size_t e = 0; e < gi.size(); ++e) {
            long long bits = gi[e];
            double key;
            std::memcpy(&key, &bits, sizeof(key));
            ga[e] = make_double2(wx[e], wy[e]);
            gb[e] = make_double2(wz[e], key);
        }
        int rc;
        if ((rc = upload(c, gs, &d.gslice)) || (rc = upload(c, ga, &d.gwa)) || (rc = upload(c, gb, &d.gwb))) return rc;
    }
    // Green-Gauss gradients (gradient_maps.cc:59-72 leaves a Green-Gauss
    // cell's map empty, and Variable::cell_gradient, variable.h:193-216, then
    // sums face values): the cells without a map, and the face-average
    // weights of every face
    {
        std::vector<int> ggc(N, 0);
        bool any = false;
        for (int i = 0; i < N; ++i) {
            const int r = g.iperm[i];
            if (g.gptr[r + 1] == g.gptr[r]) { ggc[i] = 1; any = true; }
        }
        if (any) {
            std::vector<double2> ggw(F);
            for (int i = 0; i < F; ++i) {
                const int f = g.ifperm[i];
                if (g.aptr[f + 1] - g.aptr[f] != 2 || g.aidx[g.aptr[f]] != g.fparent[f] || g.aidx[g.aptr[f] + 1] != g.fneigh[f])
                    return fail(c, FCFD_ENOTSUP, "Green-Gauss gradients on the device need two-cell face-average stencils");
                ggw[i] = make_double2(g.aw[g.aptr[f]], g.aw[g.aptr[f] + 1]);
            }
            int rc;
            if ((rc = upload(c, ggc, &d.ggc)) || (rc = upload(c, ggw, &d.ggw))) return rc;
        }
    }
    // half-faces and flux blocks
    {
        std::vector<int> hp, hf;
        half_faces(g, hp, hf);
        d.nhf = (int)hf.size();
        std::vector<int> hn(hf.size());
        for (int i = 0; i < N; ++i) {
            if (hp[i + 1] - hp[i] > kMaxCellFaces || hp[i + 1] == hp[i]) return fail(c, FCFD_EMESH, "cell with no or more than 6 faces");
            for (int h = hp[i]; h < hp[i + 1]; ++h) {
                const int fs = hf[h], fr = g.ifperm[fs >= 0 ? fs : ~fs];
                hn[h] = g.perm[fs >= 0 ? g.fneigh[fr] : g.fparent[fr]];
            }
        }
        int rc;
        if ((rc = upload(c, hp, &d.hf_ptr)) || (rc = upload(c, hf, &d.hf_face)) || (rc = upload(c, hn, &d.hf_nb))) return rc;
    }
    // faces
    {
        std::vector<int> par(F), nei(F), bc(F);
        std::vector<double> nx(F), ny(F), nz(F), tx(F), ty(F), tz(F), ar(F), cx(F), cy(F), cz(F), wp(F, 0.5), wn(F, 0.5), wd(F), ds(F);
        std::vector<int> wslot(F, -1), wptr(1, 0), widx;   // wider face-average stencils
        std::vector<double> ww;
        for (int i = 0; i < F; ++i) {
            int f = g.ifperm[i];
            par[i] = g.perm[g.fparent[f]];
            nei[i] = g.perm[g.fneigh[f]];
            bc[i] = g.fbc[f];
            nx[i] = g.fn[3 * f]; ny[i] = g.fn[3 * f + 1]; nz[i] = g.fn[3 * f + 2];
            tx[i] = g.ft1[3 * f]; ty[i] = g.ft1[3 * f + 1]; tz[i] = g.ft1[3 * f + 2];
            ar[i] = g.farea[f];
            cx[i] = g.fc[3 * f]; cy[i] = g.fc[3 * f + 1]; cz[i] = g.fc[3 * f + 2];
            wd[i] = g.fwd[f]; ds[i] = g.fdiss[f];
            if (g.fbc[f] < 0) {
                if (g.aptr[f + 1] - g.aptr[f] == 2 && g.aidx[g.aptr[f]] == g.fparent[f] && g.aidx[g.aptr[f] + 1] == g.fneigh[f]) {
                    wp[i] = g.aw[g.aptr[f]];
                    wn[i] = g.aw[g.aptr[f] + 1];
                } else {
                    wslot[i] = (int)wptr.size() - 1;
                    for (int k = g.aptr[f]; k < g.aptr[f + 1]; ++k) {
                        widx.push_back(g.perm[g.aidx[k]]);
                        ww.push_back(g.aw[k]);
                    }
                    wptr.push_back((int)widx.size());
                }
            }
        }
        d.wslot = nullptr; d.nwide = 0; d.wptr = nullptr; d.widx = nullptr; d.ww = nullptr;
        d.wgrad = nullptr; d.wgrad_d = nullptr; d.wgrad_f = nullptr;
        d.wtptr = nullptr; d.wtface = nullptr; d.wtw = nullptr;
        const int nwide = (int)wptr.size() - 1;
        if (nwide > 0 && c->P.viscous) {   // only the viscous face gradients read the face averages
            if (c->P.kw)
                return fail(c, FCFD_ENOTSUP, "the segregated two-equation RANS model on the device needs two-cell face-average stencils");
            const size_t rows = 3 * ((size_t)d.nv - 1);
            int rc;
            if ((rc = upload(c, wslot, &d.wslot)) || (rc = upload(c, wptr, &d.wptr)) || (rc = upload(c, widx, &d.widx)) ||
                (rc = upload(c, ww, &d.ww)) || (rc = dalloc(c, rows * nwide, &d.wgrad)) || (rc = dalloc(c, rows * nwide, &d.wgrad_d)) ||
                (rc = dalloc(c, rows * nwide, &d.wgrad_f)))
                return rc;
            d.nwide = nwide;
            // transposed: per interior cell, its wide faces in ascending internal
            // face order.  A stencil may hold a boundary ghost, whose gradient is
            // its parent's copy: that entry belongs to the parent.
            auto owner = [&](int j) { return j < N ? j : g.perm[g.fparent[g.ghost_face[g.iperm[j] - N]]]; };
            std::vector<int> wtptr(N + 1, 0), wtface(widx.size());
            std::vector<double> wtw(widx.size());
            for (size_t k = 0; k < widx.size(); ++k) wtptr[owner(widx[k]) + 1]++;
            for (int i = 0; i < N; ++i) wtptr[i + 1] += wtptr[i];
            std::vector<int> next(wtptr.begin(), wtptr.end() - 1);
            for (int i = 0; i < F; ++i) {
                const int sl = wslot[i];
                if (sl < 0) continue;
                for (int k = wptr[sl]; k < wptr[sl + 1]; ++k) {
                    const int q = next[owner(widx[k])]++;
                    wtface[q] = i;
                    wtw[q] = ww[k];
                }
            }
            if ((rc = upload(c, wtptr, &d.wtptr)) || (rc = upload(c, wtface, &d.wtface)) || (rc = upload(c, wtw, &d.wtw))) return rc;
        }
        std::vector<int> gh(B);
        for (int i = 0; i < B; ++i) gh[i] = g.fperm[g.ghost_face[g.iperm[N + i] - N]];
        int rc;
        if ((rc = upload(c, par, &d.fpar)) || (rc = upload(c, nei, &d.fnei)) || (rc = upload(c, bc, &d.fbc)) ||
            (rc = upload(c, nx, &d.fnx)) || (rc = upload(c, ny, &d.fny)) || (rc = upload(c, nz, &d.fnz)) ||
            (rc = upload(c, tx, &d.t1x)) || (rc = upload(c, ty, &d.t1y)) || (rc = upload(c, tz, &d.t1z)) ||
            (rc = upload(c, ar, &d.farea)) || (rc = upload(c, cx, &d.fcx)) || (rc = upload(c, cy, &d.fcy)) ||
            (rc = upload(c, cz, &d.fcz)) || (rc = upload(c, wp, &d.fwp)) || (rc = upload(c, wn, &d.fwn)) ||
            (rc = upload(c, wd, &d.fwd)) || (rc = upload(c, ds, &d.fdiss)) || (rc = upload(c, gh, &d.gh_face)))
            return rc;
        std::vector<int4> cell(F);
        std::vector<double2> geo((size_t)6 * F);
        for (int i = 0; i < F; ++i) {
            cell[i] = make_int4(par[i], nei[i], bc[i], 0);
            geo[i] = make_double2(nx[i], ny[i]);
            geo[(size_t)F + i] = make_double2(nz[i], tx[i]);
            geo[(size_t)2 * F + i] = make_double2(ty[i], tz[i]);
            geo[(size_t)3 * F + i] = make_double2(ar[i], cx[i]);
            geo[(size_t)4 * F + i] = make_double2(cy[i], cz[i]);
            geo[(size_t)5 * F + i] = make_double2(wp[i], wn[i]);
        }
        if ((rc = upload(c, cell, &d.fcell)) || (rc = upload(c, geo, &d.fgeo))) return rc;
    }
    {
        int rc;
        if ((rc = upload(c, c->bch, &d.bc))) return rc;
        const size_t nv = d.nv;
        const size_t grows = 2 * pair_groups(3 * (int)nv), prows = 2 * pair_groups((int)nv);   // row pairs
        if ((rc = dalloc(c, grows * ld, &c->G)) || (rc = dalloc(c, prows * ld, &c->PHI)) ||
            (rc = dalloc(c, (size_t)kDrefSize, &c->dref)) || (rc = dalloc(c, 4, &c->status)) ||
            (rc = dalloc(c, nv * g.F, &c->FX)))
            return rc;
        if (nv == 6 && (rc = dalloc(c, ld, &c->SRC))) return rc;
    }
    if (c->gpu_wall_distance) {   // brute force on the device, then mirrored into the host metrics
        std::vector<double> wx, wy, wz, wnx, wny, wnz;
        std::vector<int> is_wall(F, 0);
        for (int f = 0; f < F; ++f) {   // reference face order (first closest face wins ties)
            int b = g.fbc[f];
            if (b >= 0 && c->opt.bc[b].type == FCFD_BC_WALL && c->opt.bc[b].kind != FCFD_KIND_SLIP) {
                wx.push_back(g.fc[3 * f]); wy.push_back(g.fc[3 * f + 1]); wz.push_back(g.fc[3 * f + 2]);
                wnx.push_back(g.fn[3 * f]); wny.push_back(g.fn[3 * f + 1]); wnz.push_back(g.fn[3 * f + 2]);
                is_wall[g.fperm[f]] = 1;
            }
        }
        const int nw = (int)wx.size();
        const double *dwx, *dwy, *dwz, *dwnx, *dwny, *dwnz;
        const int* dis;
        int rc;
        if ((rc = upload(c, wx, &dwx)) || (rc = upload(c, wy, &dwy)) || (rc = upload(c, wz, &dwz)) || (rc = upload(c, wnx, &dwnx)) ||
            (rc = upload(c, wny, &dwny)) || (rc = upload(c, wnz, &dwnz)) || (rc = upload(c, is_wall, &dis)))
            return rc;
        launch_wall_distance(T, d.ccx, d.ccy, d.ccz, nw, dwx, dwy, dwz, const_cast<double*>(d.wd), nullptr);
        // face distances feed only the Steger-Warming dissipation factor
        const bool faces = c->opt.convective_flux == FCFD_FLUX_SW;
        if (faces)
            launch_face_wall(F, d.fcx, d.fcy, d.fcz, d.fnx, d.fny, d.fnz, dis, nw, dwx, dwy, dwz, dwnx, dwny, dwnz,
                             const_cast<double*>(d.fwd), const_cast<double*>(d.fdiss), nullptr);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipDeviceSynchronize());
        std::vector<double> tw(ld), tf(F), td(F);
        HIPCHK(c, hipMemcpy(tw.data(), d.wd, ld * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(tf.data(), d.fwd, F * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(td.data(), d.fdiss, F * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < T; ++i) g.wdist[g.iperm[i]] = tw[i];
        if (faces)
            for (int i = 0; i < F; ++i) { g.fwd[g.ifperm[i]] = tf[i]; g.fdiss[g.ifperm[i]] = td[i]; }
    }
    if (c->P.kw) {   // the two-equation model's fields (mu_t read by every VISC = kViscRans face kernel)
        int rc;
        KwDev& k = c->kw;
        KwTerms& t = c->kwt;
        const size_t b2 = 2 * (size_t)pair_groups(4);
        if ((rc = dalloc(c, 2 * (size_t)ld, &k.K)) || (rc = dalloc(c, (size_t)ld, &k.mut)) || (rc = dalloc(c, 6 * (size_t)ld, &k.GK)) ||
            (rc = dalloc(c, (size_t)ld, &k.strain)) || (rc = dalloc(c, 6 * (size_t)F, &t.FV)) || (rc = dalloc(c, b2 * F, &t.JL)) ||
            (rc = dalloc(c, b2 * F, &t.JR)) || (rc = dalloc(c, 2 * (size_t)ld, &t.rhs)) || (rc = dalloc(c, b2 * ld, &t.D)) ||
            (rc = dalloc(c, 1, &t.lam_last)) || (rc = dalloc(c, 2 * (size_t)ld, &c->kw_dU)) || (rc = dalloc(c, 3 * 256, &c->kw_part)) ||
            (rc = dalloc(c, 4, &c->kw_out)) || (rc = dalloc(c, 3 * (size_t)std::max(c->opt.n_bc, 1), &c->kw_inl)) ||
            (rc = dalloc(c, b2 * ld, &c->kw_Dinv)))
            return rc;
        d.kw_mut = k.mut;
        std::vector<int> f0(std::max(c->opt.n_bc, 1), -1);
        for (int f = g.F - 1; f >= 0; --f)
            if (g.fbc[f] >= 0) f0[g.fbc[f]] = g.fperm[f];
        if ((rc = upload(c, f0, &c->kw_f0))) return rc;
    }
    c->device = true;
    return FCFD_OK;
}

thread_local std::string g_create_err;

int fold_slot(fcfd_ctx* c, int slot) {
    hipEvent_t* e = &c->ev[(fcfd_ctx::kStages + 1) * slot];
    HIPCHK(c, hipEventSynchronize(e[fcfd_ctx::kStages]));
    for (int k = 0; k < fcfd_ctx::kStages; ++k) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, e[k], e[k + 1]));
        c->stage_ms[k] += ms;
    }
    c->stage_calls++;
    return FCFD_OK;
}

int create_impl(const fcfd_mesh* mesh, const fcfd_options* opt, fcfd_ctx** out, bool device) {
    if (!out) return FCFD_EINVAL;
    *out = nullptr;
    fcfd_ctx* c = new (std::nothrow) fcfd_ctx();
    if (!c) return FCFD_ENOMEM;
    if (!mesh || !opt) { g_create_err = "null mesh or options"; delete c; return FCFD_EINVAL; }
    c->opt = *opt;
    int rc = validate_options(c);
    if (!rc) {
        long long nw = 0;
        for (int b = 0; b < opt->n_bc; ++b) nw += (opt->bc[b].type == FCFD_BC_WALL && opt->bc[b].kind != FCFD_KIND_SLIP);
        // brute force on the host for small meshes; the device computes them
        // for large ones (the same exact minimum, set_bcs.cc:178-249)
        // (a host-only context of a large mesh leaves them unset: 1e20)
        long long host_max = 2000000LL;   // FCFD_HOST_WALL_DISTANCE_MAX_CELLS overrides (tests force the device path)
        if (const char* e = std::getenv("FCFD_HOST_WALL_DISTANCE_MAX_CELLS")) host_max = std::atoll(e);
        bool host_wd = nw == 0 || (long long)mesh->n_cells < host_max;
        rc = build_host_grid(mesh, opt, c->hg, c->err, host_wd);
        c->gpu_wall_distance = !host_wd && device;
    }
    if (!rc) {
        // internal cell order: space-filling curve (default) or reverse
        // Cuthill-McKee (FCFD_CELL_ORDER=rcm); results do not depend on it
        int order = kOrderSfc;
        if (const char* e = std::getenv("FCFD_CELL_ORDER")) order = std::strcmp(e, "rcm") == 0 ? kOrderRcm : kOrderSfc;
        renumber(c->hg, order);
        set_params(c);
        if (device) rc = build_device(c);
    }
    if (rc) {
        g_create_err = c->err;
        for (void* p : c->allocs) (void)hipFree(p);
        delete c;
        return rc;
    }
    *out = c;
    return FCFD_OK;
}

}  // namespace

extern "C" {

int fcfd_create(const fcfd_mesh* mesh, const fcfd_options* opt, fcfd_ctx** out) { return create_impl(mesh, opt, out, true); }
int fcfd_create_host(const fcfd_mesh* mesh, const fcfd_options* opt, fcfd_ctx** out) { return create_impl(mesh, opt, out, false); }

void fcfd_destroy(fcfd_ctx* c) {
    if (!c) return;
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    if (c->mg_exec) (void)hipGraphExecDestroy(c->mg_exec);
    if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
    for (void* p : c->allocs) (void)hipFree(p);
    mg_free(c->mg);
    delete c;
}

const char* fcfd_last_error(const fcfd_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int fcfd_get_info(const fcfd_ctx* c, fcfd_info* info) {
    if (!c || !info) return FCFD_EINVAL;
    const HostGrid& g = c->hg;
    std::memset(info, 0, sizeof *info);
    info->n_nodes = g.nnode; info->n_cells = g.N; info->n_ghosts = g.B; info->n_faces = g.F; info->n_bfaces = g.B;
    info->ld_cell = ((g.N + g.B + 63) / 64) * 64;
    info->ld_face = g.F;
    info->n_gradmap = (int)g.gidx.size();
    info->n_faceavg = (int)g.aidx.size();
    info->n_nbr = (int)g.nb.size();
    info->n_halffaces = (int)g.cfaces.size();
    info->nv = (c->opt.turbulence == FCFD_TURB_SA) ? 6 : 5;
    info->n_nodeavg = (int)g.naidx.size();
    return FCFD_OK;
}

int fcfd_metrics(const fcfd_ctx* c, fcfd_metrics_out* m) {
    if (!c || !m) return FCFD_EINVAL;
    const HostGrid& g = c->hg;
    const int F = g.F, N = g.N, T = g.N + g.B;
    auto cp_i = [](int* dst, const std::vector<int>& src) { if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(int)); };
    auto cp_d = [](double* dst, const std::vector<double>& src, size_t n) { if (dst && n) std::memcpy(dst, src.data(), n * sizeof(double)); };
    cp_i(m->face_parent, g.fparent); cp_i(m->face_neighbor, g.fneigh); cp_i(m->face_bc, g.fbc);
    cp_i(m->face_node_ptr, g.fnptr); cp_i(m->face_nodes, g.fnodes);
    cp_d(m->face_centroid, g.fc, 3 * (size_t)F); cp_d(m->face_normal, g.fn, 3 * (size_t)F);
    cp_d(m->face_area, g.farea, F); cp_d(m->face_tangent1, g.ft1, 3 * (size_t)F);
    cp_d(m->cell_centroid, g.cc, 3 * (size_t)T); cp_d(m->cell_volume, g.vol, T);
    cp_d(m->cell_lengthscale, g.lscale, N); cp_d(m->wall_distance, g.wdist, T);
    cp_i(m->cell_face_ptr, g.cfptr); cp_i(m->cell_faces, g.cfaces);
    cp_i(m->nbr_ptr, g.nbptr); cp_i(m->nbr, g.nb);
    cp_i(m->grad_ptr, g.gptr); cp_i(m->grad_idx, g.gidx); cp_d(m->grad_w, g.gw, g.gw.size());
    cp_i(m->avg_ptr, g.aptr); cp_i(m->avg_idx, g.aidx); cp_d(m->avg_w, g.aw, g.aw.size());
    cp_i(m->cell_perm, g.perm);
    if (m->grid_lengthscale) m->grid_lengthscale[0] = g.glscale;
    cp_d(m->node_xyz, g.node, 3 * (size_t)g.nnode);
    cp_i(m->node_avg_ptr, g.naptr); cp_i(m->node_avg_idx, g.naidx); cp_d(m->node_avg_w, g.naw, g.naw.size());
    return FCFD_OK;
}

int fcfd_state_to_device(fcfd_ctx* c, const double* W_host, double* W_dev) {
    if (!c || !W_host || !W_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data (created with fcfd_create_host)");
    const int N = c->hg.N, ld = c->dg.ld;
    const int nv = c->P.nv;
    std::vector<double> soa((size_t)nv * ld, 0.0);
    for (int i = 0; i < N; ++i) {
        int r = c->hg.iperm[i];
        for (int v = 0; v < nv; ++v) soa[(size_t)v * ld + i] = W_host[(size_t)nv * r + v];
    }
    HIPCHK(c, hipMemcpy(W_dev, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
    return FCFD_OK;
}

int fcfd_cells_to_host(fcfd_ctx* c, const double* X_dev, int nvar, int include_ghosts, double* X_host) {
    if (!c || !X_dev || !X_host || nvar <= 0) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    const int N = c->hg.N, B = c->hg.B, ld = c->dg.ld;
    const int n = include_ghosts ? N + B : N;
    std::vector<double> soa((size_t)nvar * ld);
    HIPCHK(c, hipMemcpy(soa.data(), X_dev, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        int r = c->hg.iperm[i];
        for (int v = 0; v < nvar; ++v) X_host[(size_t)nvar * r + v] = soa[(size_t)v * ld + i];
    }
    return FCFD_OK;
}

int fcfd_residual(fcfd_ctx* c, double* W_dev, double* R_dev, void* stream) {
    if (!c || !W_dev || !R_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t* e = nullptr;
    if (c->profiling) {
        if (c->ring_count == fcfd_ctx::kRing) {   // fold the oldest slot (long finished)
            int old = (c->ring_head + fcfd_ctx::kRing - c->ring_count) % fcfd_ctx::kRing;
            int rc = fold_slot(c, old);
            if (rc) return rc;
            c->ring_count--;
        }
        e = &c->ev[(fcfd_ctx::kStages + 1) * c->ring_head];
        HIPCHK(c, hipEventRecord(e[0], s));
    }
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);
    if (e) HIPCHK(c, hipEventRecord(e[1], s));
    launch_grad_lim(c->dg, c->P, W_dev, c->G, c->PHI, c->dref, s, c->SRC);
    if (e) HIPCHK(c, hipEventRecord(e[2], s));
    launch_flux(c->dg, c->P, W_dev, c->G, c->PHI, c->FX, R_dev, c->status, s);
    if (e) HIPCHK(c, hipEventRecord(e[3], s));
    launch_residual_sum(c->dg, c->SRC, c->FX, R_dev, s);
    if (e) {
        HIPCHK(c, hipEventRecord(e[4], s));
        c->ring_head = (c->ring_head + 1) % fcfd_ctx::kRing;
        c->ring_count++;
    }
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

int fcfd_profile(fcfd_ctx* c, int enable, double* stage_ms, long long* calls) {
    if (!c || !c->device) return FCFD_EINVAL;
    while (c->ring_count > 0) {
        int old = (c->ring_head + fcfd_ctx::kRing - c->ring_count) % fcfd_ctx::kRing;
        int rc = fold_slot(c, old);
        if (rc) return rc;
        c->ring_count--;
    }
    if (stage_ms) for (int k = 0; k < fcfd_ctx::kStages; ++k) stage_ms[k] = c->stage_ms[k];
    if (calls) *calls = c->stage_calls;
    if (enable >= 0) {
        if (enable && c->ev.empty()) {
            c->ev.resize((fcfd_ctx::kStages + 1) * fcfd_ctx::kRing, nullptr);
            for (auto& ev : c->ev) HIPCHK(c, hipEventCreate(&ev));
        }
        c->profiling = enable != 0;
        for (int k = 0; k < fcfd_ctx::kStages; ++k) c->stage_ms[k] = 0.0;
        c->stage_calls = 0;
        c->ring_head = 0;
    }
    return FCFD_OK;
}

int fcfd_gradients_limiters(fcfd_ctx* c, double* W_dev, double* G_dev, double* PHI_dev, void* stream) {
    if (!c || !W_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    hipStream_t s = (hipStream_t)stream;
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);
    launch_grad_lim(c->dg, c->P, W_dev, G_dev ? G_dev : c->G, PHI_dev ? PHI_dev : c->PHI, c->dref, s, nullptr);
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

int fcfd_internal_buffers(fcfd_ctx* c, const double** grad, const double** phi) {
    if (!c || !c->device) return FCFD_EINVAL;
    if (grad) *grad = c->G;
    if (phi) *phi = c->PHI;
    return FCFD_OK;
}

int fcfd_gradients_to_host(fcfd_ctx* c, double* G_host, double* PHI_host) {
    if (!c || !c->device) return FCFD_EINVAL;
    const int N = c->hg.N, B = c->hg.B, ld = c->dg.ld, nv = c->dg.nv;
    std::vector<double> g((size_t)2 * pair_groups(3 * nv) * ld), p((size_t)2 * pair_groups(nv) * ld);
    HIPCHK(c, hipDeviceSynchronize());
    HIPCHK(c, hipMemcpy(g.data(), c->G, g.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(p.data(), c->PHI, p.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < N + B; ++i) {
        const int r = c->hg.iperm[i];
        if (G_host)
            for (int k = 0; k < 3 * nv; ++k) G_host[(size_t)3 * nv * r + k] = g[pidx(ld, k, i)];
        if (PHI_host && i < N)
            for (int v = 0; v < nv; ++v) PHI_host[(size_t)nv * r + v] = p[pidx(ld, v, i)];
    }
    return FCFD_OK;
}

int fcfd_status(fcfd_ctx* c, int clear) {
    if (!c || !c->device) return FCFD_EINVAL;
    int h = 0;
    HIPCHK(c, hipMemcpy(&h, c->status, sizeof(int), hipMemcpyDeviceToHost));
    if (clear) HIPCHK(c, hipMemset(c->status, 0, sizeof(int)));
    if (h & 1) return fail(c, FCFD_EBC, "an inlet became supersonic; the thermodynamic state must be specified (ns_apply_bcs.cc:54-59)");
    return FCFD_OK;
}

static int check_dt(fcfd_ctx* c, int dt_type, double value) {
    if (dt_type != FCFD_DT_FIXED && dt_type != FCFD_DT_CFL_MAX && dt_type != FCFD_DT_CFL_LOCAL)
        return fail(c, FCFD_EINVAL, "unknown time step type");
    if (!(value > 0.0)) return fail(c, FCFD_EINVAL, "the step size or CFL number must be positive");
    if (!c->dt_red) return dalloc(c, 2 + 256, &c->dt_red);
    return FCFD_OK;
}

int fcfd_time_step(fcfd_ctx* c, const double* W_dev, int dt_type, double cfl, double* dt_dev, double* max_cfl, void* stream) {
    if (!c || !W_dev || !dt_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    if (int rc = check_dt(c, dt_type, cfl)) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_time_step(c->dg, c->P, W_dev, dt_type, cfl, dt_dev, c->dt_red, s);
    HIPCHK(c, hipGetLastError());
    if (max_cfl) {
        if (dt_type == FCFD_DT_CFL_LOCAL) {
            *max_cfl = cfl;   // time_step.cc:207-210
        } else {
            launch_max_cfl(c->dg, c->P, W_dev, dt_type == FCFD_DT_CFL_MAX ? c->dt_red : nullptr, cfl, c->dt_red, s);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(max_cfl, c->dt_red + 1, sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(c, hipStreamSynchronize(s));
        }
    }
    return FCFD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scratch of fcfd_time_step_noalloc: the CFL_MAX minimum (slot 0), the
// largest CFL number (slot 1) and the 256 block partials of the two-pass
// reductions (slots 2..257), the same layout as the context's lazily
// allocated dt_red.  Its size does not depend on the mesh.  In host code the
// array's name is no device address: kernels get the per-device address
// from hipGetSymbolAddress, looked up once per device with that device
// current.  Every call on one GPU shares that GPU's array.
// ---------------------------------------------------------------------------
namespace {
constexpr int kDtScratch = 2 + 256;
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[kDtScratch];
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(fcfd_ctx* c, double** out) {
    int dev = -1;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(c, FCFD_ENOTSUP, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return FCFD_OK;
}
}  // namespace

extern "C" {

int fcfd_time_step_noalloc(fcfd_ctx* c, const double* W_dev, int dt_type, double cfl, double* dt_dev, double* max_cfl_dev,
                           void* stream) {
    if (!c || !W_dev || !dt_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    if (dt_type != FCFD_DT_FIXED && dt_type != FCFD_DT_CFL_MAX && dt_type != FCFD_DT_CFL_LOCAL)
        return fail(c, FCFD_EINVAL, "unknown time step type");
    if (!(cfl > 0.0)) return fail(c, FCFD_EINVAL, "the step size or CFL number must be positive");
    double* red = nullptr;
    if (int rc = scratch_for_current_device(c, &red)) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_time_step(c->dg, c->P, W_dev, dt_type, cfl, dt_dev, red, s);
    if (max_cfl_dev) {
        if (dt_type == FCFD_DT_CFL_LOCAL) {
            launch_store_scalar(nullptr, cfl, max_cfl_dev, s);   // time_step.cc:207-210
        } else {
            launch_max_cfl(c->dg, c->P, W_dev, dt_type == FCFD_DT_CFL_MAX ? red : nullptr, cfl, red, s);
            launch_store_scalar(red + 1, 0.0, max_cfl_dev, s);
        }
    }
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

int fcfd_fd_jacobian(fcfd_ctx* c, double* W_dev, const double* dWprev_dev, const double* dt_dev, void* stream) {
    if (!c || !W_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    const DevGrid& g = c->dg;
    if (!c->JL) {
        int rc;
        const size_t b2 = 2 * (size_t)pair_groups(g.nv * g.nv);   // block entries in row pairs
        if ((rc = dalloc(c, b2 * g.F, &c->JL)) || (rc = dalloc(c, b2 * g.F, &c->JR)) ||
            (rc = dalloc(c, b2 * g.ld, &c->Dg)) || (rc = dalloc(c, (size_t)g.nv * g.ld, &c->zero_upd)))
            return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    launch_ghosts(g, c->P, W_dev, c->status, s);
    launch_grad_lim(g, c->P, W_dev, c->G, c->PHI, c->dref, s);
    launch_fd_jacobian(g, c->P, W_dev, c->G, c->PHI, dWprev_dev ? dWprev_dev : c->zero_upd, dt_dev, c->JL, c->JR, c->Dg,
                       c->status, s);
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

int fcfd_jacobian_buffers(fcfd_ctx* c, const double** JL, const double** JR, const double** D) {
    if (!c || !c->device || !c->JL) return fail(c, FCFD_EINVAL, "no Jacobian has been assembled");
    if (JL) *JL = c->JL;
    if (JR) *JR = c->JR;
    if (D) *D = c->Dg;
    return FCFD_OK;
}

int fcfd_jacobian_to_host(fcfd_ctx* c, double* JL_host, double* JR_host, double* D_host) {
    if (!c || !c->device || !c->JL) return fail(c, FCFD_EINVAL, "no Jacobian has been assembled");
    HIPCHK(c, hipDeviceSynchronize());
    const HostGrid& hg = c->hg;
    const int F = hg.F, N = hg.N, ld = c->dg.ld;
    const int b2 = c->dg.nv * c->dg.nv;
    const size_t rows = 2 * (size_t)pair_groups(b2);   // stored in row pairs
    std::vector<double> tmp(rows * std::max(F, ld));
    double* outs[2] = {JL_host, JR_host};
    const double* srcs[2] = {c->JL, c->JR};
    for (int k = 0; k < 2; ++k) {
        if (!outs[k]) continue;
        HIPCHK(c, hipMemcpy(tmp.data(), srcs[k], rows * F * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < F; ++i) {
            int f = hg.ifperm[i];
            for (int e = 0; e < b2; ++e) outs[k][(size_t)b2 * f + e] = tmp[pidx((size_t)F, e, i)];
        }
    }
    if (D_host) {
        HIPCHK(c, hipMemcpy(tmp.data(), c->Dg, rows * ld * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < N; ++i) {
            int r = hg.iperm[i];
            for (int e = 0; e < b2; ++e) D_host[(size_t)b2 * r + e] = tmp[pidx((size_t)ld, e, i)];
        }
    }
    return FCFD_OK;
}

static int ensure_tangent(fcfd_ctx* c) {
    if (c->dG) return FCFD_OK;
    const DevGrid& g = c->dg;
    int rc;
    const size_t nv = g.nv;
    if ((rc = dalloc(c, 2 * pair_groups(3 * (int)nv) * g.ld, &c->dG)) || (rc = dalloc(c, 2 * pair_groups((int)nv) * g.ld, &c->dPHI)) ||
        (rc = dalloc(c, nv * g.F, &c->dFX)) || (rc = dalloc(c, nv * g.ld, &c->jv_in)) ||
        (rc = dalloc(c, nv * g.ld, &c->jv_out)) || (rc = dalloc(c, nv * g.ld, &c->jv_R)) ||
        (nv == 6 && (rc = dalloc(c, g.ld, &c->dSRC))))
        return rc;
    return FCFD_OK;
}

// Sliced-ELL index of the off-diagonal blocks (64-cell slices): per cell
// its half-faces whose other cell is interior, in half-face (reference face)
// order, and the block storage of the precision the Krylov solver needs.
static int ensure_sell(fcfd_ctx* c, bool fp32) {
    const DevGrid& g = c->dg;
    int rc;
    if (!c->sell.soff) {
        const HostGrid& hg = c->hg;
        const int N = g.N, ns = (N + 63) / 64;
        std::vector<int> hp, hf;
        half_faces(hg, hp, hf);
        std::vector<int> soff(ns + 1, 0), sface, snb;
        for (int sl = 0; sl < ns; ++sl) {
            int width = 0;
            for (int lane = 0; lane < 64; ++lane) {
                const int i = 64 * sl + lane;
                if (i >= N) break;
                int k = 0;
                for (int h = hp[i]; h < hp[i + 1]; ++h) {
                    const int fs = hf[h], fr = hg.ifperm[fs >= 0 ? fs : ~fs];
                    if (hg.perm[fs >= 0 ? hg.fneigh[fr] : hg.fparent[fr]] < N) ++k;
                }
                width = std::max(width, k);
            }
            const size_t base = sface.size();
            sface.resize(base + (size_t)64 * width, kSellPad);
            snb.resize(base + (size_t)64 * width, -1);
            for (int lane = 0; lane < 64; ++lane) {
                const int i = 64 * sl + lane;
                if (i >= N) break;
                int k = 0;
                for (int h = hp[i]; h < hp[i + 1]; ++h) {
                    const int fs = hf[h], fr = hg.ifperm[fs >= 0 ? fs : ~fs];
                    const int o = hg.perm[fs >= 0 ? hg.fneigh[fr] : hg.fparent[fr]];
                    if (o >= N) continue;
                    sface[base + (size_t)64 * k + lane] = fs;
                    snb[base + (size_t)64 * k + lane] = o;
                    ++k;
                }
            }
            soff[sl + 1] = soff[sl] + width;
        }
        c->sell.nslice = ns;
        c->sell.nslot = (size_t)soff[ns];
        if ((rc = upload(c, soff, &c->sell.soff)) || (rc = upload(c, sface, &c->sell.sface)) || (rc = upload(c, snb, &c->sell.snb)))
            return rc;
    }
    const size_t n = c->sell.nslot * 64 * (size_t)(2 * pair_groups(g.nv * g.nv));
    if (fp32 && !c->sell32 && (rc = dalloc(c, n, &c->sell32))) return rc;
    if (!fp32 && !c->sell64 && (rc = dalloc(c, n, &c->sell64))) return rc;
    return FCFD_OK;
}

static int ensure_tangent32(fcfd_ctx* c) {
    if (c->dG32) return FCFD_OK;
    const DevGrid& g = c->dg;
    int rc;
    const size_t nv = g.nv;
    if ((rc = dalloc(c, 2 * pair_groups(3 * (int)nv) * g.ld, &c->dG32)) || (rc = dalloc(c, 2 * pair_groups((int)nv) * g.ld, &c->dPHI32)) ||
        (rc = dalloc(c, nv * g.F, &c->dFX32)) || (rc = dalloc(c, nv * g.ld, &c->jv_in32)) || (rc = dalloc(c, nv * g.ld, &c->jv_R32)) ||
        (nv == 6 && (rc = dalloc(c, g.ld, &c->dSRC32))))
        return rc;
    if (c->P.order != FCFD_ORDER_FIRST && c->P.limiter != FCFD_LIMITER_NONE &&
        ((rc = dalloc(c, nv * g.ld, &c->lin32.nb)) || (rc = dalloc(c, nv * g.ld, &c->lin32.a)) ||
         (rc = dalloc(c, 3 * nv * g.ld, &c->lin32.gv))))
        return rc;
    return FCFD_OK;
}

// Reverse-sweep connectivity (half-face twins, ghost half-faces, transposed
// gradient map), the per-BC load face lists and the adjoint workspace.
static int ensure_adjoint(fcfd_ctx* c) {
    if (c->adj_ready) return FCFD_OK;
    const HostGrid& g = c->hg;
    const DevGrid& d = c->dg;
    const int N = g.N, B = g.B, F = g.F, T = N + B;
    int rc;
    std::vector<int> hp, hf;
    half_faces(g, hp, hf);
    const int nhf = (int)hf.size();
    {
        std::vector<int> par_h(F, -1), nei_h(F, -1);
        for (int h = 0; h < nhf; ++h) {
            if (hf[h] >= 0) par_h[hf[h]] = h;
            else nei_h[~hf[h]] = h;
        }
        std::vector<int> twin(nhf), ghh(B);
        for (int h = 0; h < nhf; ++h) twin[h] = hf[h] >= 0 ? nei_h[hf[h]] : par_h[~hf[h]];
        for (int i = 0; i < B; ++i) ghh[i] = par_h[g.fperm[g.ghost_face[g.iperm[N + i] - N]]];
        if ((rc = upload(c, twin, &c->ag.hf_twin)) || (rc = upload(c, ghh, &c->ag.gh_hf))) return rc;
    }
    {
        // Transposed gradient stencils.  A mapped cell's gradient is sum_k w_k W_k
        // over its map; a Green-Gauss cell's is sum_faces (n A / V) times the
        // face value, w_p W_parent + w_n W_neighbour (or ghost), except for the
        // variables the face's BC holds fixed, whose face value is a constant.
        // Each cell visits its entries (j, 3 weights, fixed-variable bits).
        const bool gg = d.ggc != nullptr;
        auto visit = [&](int i, auto&& emit) {
            const int r = g.iperm[i];
            for (int e = g.gptr[r]; e < g.gptr[r + 1]; ++e) emit(g.perm[g.gidx[e]], g.gw[3 * e], g.gw[3 * e + 1], g.gw[3 * e + 2], 0);
            if (!gg || g.gptr[r + 1] != g.gptr[r]) return;
            for (int k = g.cfptr[r]; k < g.cfptr[r + 1]; ++k) {
                const int f = g.cfaces[k];
                const double sgn = g.fparent[f] == r ? 1.0 : -1.0, s = sgn * g.farea[f] / g.vol[r];
                const double ax = g.fn[3 * f] * s, ay = g.fn[3 * f + 1] * s, az = g.fn[3 * f + 2] * s;
                int fixed = 0;
                if (g.fbc[f] >= 0) {
                    const BcDev& b = c->bch[g.fbc[f]];
                    fixed = (b.p_fixed ? 1 : 0) | (b.V_fixed ? 14 : 0) | (b.T_fixed ? 16 : 0);
                }
                const double wp = g.aw[g.aptr[f]], wn = g.aw[g.aptr[f] + 1];
                emit(g.perm[g.fparent[f]], ax * wp, ay * wp, az * wp, fixed);
                emit(g.perm[g.fneigh[f]], ax * wn, ay * wn, az * wn, fixed);
            }
        };
        std::vector<int> tp(T + 1, 0);
        for (int i = 0; i < N; ++i) visit(i, [&](int j, double, double, double, int) { tp[j + 1]++; });
        for (int j = 0; j < T; ++j) tp[j + 1] += tp[j];
        std::vector<int> next(tp.begin(), tp.end() - 1), tc(tp[T]), tm(gg ? tp[T] : 0);
        std::vector<double> wx(tp[T]), wy(tp[T]), wz(tp[T]);
        for (int i = 0; i < N; ++i)   // ascending internal cell: each list ends up in ascending c
            visit(i, [&](int j, double x, double y, double z, int fixed) {
                const int pos = next[j]++;
                tc[pos] = i;
                wx[pos] = x; wy[pos] = y; wz[pos] = z;
                if (gg) tm[pos] = fixed;
            });
        c->ag.tgm = nullptr;
        if ((rc = upload(c, tp, &c->ag.tg_ptr)) || (rc = upload(c, tc, &c->ag.tg_cell)) || (rc = upload(c, wx, &c->ag.tgx)) ||
            (rc = upload(c, wy, &c->ag.tgy)) || (rc = upload(c, wz, &c->ag.tgz)) || (gg && (rc = upload(c, tm, &c->ag.tgm))))
            return rc;
    }
    {
        const int nbc = c->opt.n_bc;
        std::vector<int> lp(nbc + 1, 0), lf;
        for (int b = 0; b < nbc; ++b) {
            if (c->opt.bc[b].include_in_loads)
                for (int f = 0; f < F; ++f)
                    if (g.fbc[f] == b) lf.push_back(g.fperm[f]);
            lp[b + 1] = (int)lf.size();
        }
        c->n_loadfaces = (int)lf.size();
        c->ag.n_bc = nbc;
        if ((rc = upload(c, lp, &c->ag.lf_ptr)) || (rc = upload(c, lf, &c->ag.lf_face)) ||
            (rc = dalloc(c, (size_t)6 * std::max(c->n_loadfaces, 1), &c->loads_tmp)) ||
            (rc = dalloc(c, (size_t)6 * std::max(nbc, 1), &c->loads_out)))
            return rc;
    }
    const size_t nv = d.nv;
    const size_t nz = 4 * nv + ((c->P.viscous || nv == 6) ? (nv == 6 ? 15 : 12) : 0);   // face inputs per face
    if ((rc = dalloc(c, nz * F, &c->aw.Z)) || (rc = dalloc(c, nv * d.ld, &c->aw.Wb1)) ||
        (rc = dalloc(c, 3 * nv * d.ld, &c->aw.Gb)) || (rc = dalloc(c, nv * nhf, &c->aw.LH)) ||
        (rc = dalloc(c, nv * d.ld, &c->aw.Wb2)))
        return rc;
    c->adj_ready = true;
    return FCFD_OK;
}

int fcfd_vjp(fcfd_ctx* c, double* W_dev, const double* Rbar_dev, const double* fx_weight_dev, double* Wbar_dev, void* stream) {
    if (!c || !W_dev || !Rbar_dev || !Wbar_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    int rc = ensure_adjoint(c);
    if (rc) return rc;
    launch_vjp(c->dg, c->ag, c->P, W_dev, c->G, c->PHI, c->dref, Rbar_dev, fx_weight_dev, c->aw, Wbar_dev, c->status,
               (hipStream_t)stream);
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

int fcfd_loads(fcfd_ctx* c, double* W_dev, double* loads_host, void* stream) {
    if (!c || !W_dev || !loads_host) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    int rc = ensure_adjoint(c);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);
    launch_grad_lim(c->dg, c->P, W_dev, c->G, c->PHI, c->dref, s);
    launch_loads(c->dg, c->ag, c->P, W_dev, c->G, c->PHI, c->opt.moment_center[0], c->opt.moment_center[1],
                 c->opt.moment_center[2], c->loads_tmp, c->n_loadfaces, c->loads_out, c->status, s);
    HIPCHK(c, hipGetLastError());
    if (c->opt.n_bc > 0) {
        HIPCHK(c, hipMemcpyAsync(loads_host, c->loads_out, sizeof(double) * 6 * c->opt.n_bc, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
    }
    return FCFD_OK;
}

int fcfd_convective_flux(fcfd_ctx* c, int flux, int n, const double* WL_dev, const double* WR_dev, const double* normal_dev,
                         const double* tangent1_dev, const double* area_dev, const double* diss_dev, const double* wall_dist_dev,
                         double* flux_dev, double* weightL_dev, void* stream) {
    if (!c) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    if (flux < FCFD_FLUX_ROE || flux > FCFD_FLUX_SW) return fail(c, FCFD_EINVAL, "unknown convective flux");
    if (n < 0) return fail(c, FCFD_EINVAL, "negative face count");
    if (n == 0) return FCFD_OK;
    if (!WL_dev || !WR_dev || !normal_dev || !tangent1_dev || !area_dev || !flux_dev)
        return fail(c, FCFD_EINVAL, "fcfd_convective_flux: a required array is NULL");
    launch_flux_leaf(c->P, flux, n, WL_dev, WR_dev, normal_dev, tangent1_dev, area_dev, diss_dev, wall_dist_dev, flux_dev,
                     weightL_dev, (hipStream_t)stream);
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

int fcfd_surface(fcfd_ctx* c, double* W_dev, double* mdot_host, double* qdot_host, double* tau_host, void* stream) {
    if (!c || !W_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    const HostGrid& hg = c->hg;
    const int F = hg.F;
    double* out = nullptr;
    int rc = dalloc(c, (size_t)5 * F, &out);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);
    launch_grad_lim(c->dg, c->P, W_dev, c->G, c->PHI, c->dref, s);
    launch_surface(c->dg, c->P, W_dev, c->G, c->PHI, out, c->status, s);
    std::vector<double> h((size_t)5 * F);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(c, out);
    if (e != hipSuccess) return fail(c, FCFD_EHIP, std::string("fcfd_surface: ") + hipGetErrorString(e));
    for (int i = 0; i < F; ++i) {
        const int f = hg.ifperm[i];
        if (mdot_host) mdot_host[f] = h[i];
        if (qdot_host) qdot_host[f] = h[(size_t)F + i];
        if (tau_host)
            for (int k = 0; k < 3; ++k) tau_host[3 * (size_t)f + k] = h[(size_t)(2 + k) * F + i];
    }
    return FCFD_OK;
}

// the values of one BC as a seed direction (the fields of fcfd_bc that enter
// the residual: p, T, rho, V, p_total, T_total, mdot, qdot, nut)
static BcSeed make_seed(int index, const fcfd_bc& d) {
    BcSeed q{};
    q.index = index;
    q.p = d.p; q.T = d.T; q.rho = d.rho; q.V[0] = d.V[0]; q.V[1] = d.V[1]; q.V[2] = d.V[2];
    q.p_total = d.p_total; q.T_total = d.T_total; q.mdot = d.mdot; q.qdot = d.qdot; q.nut = d.nut;
    return q;
}

static int param_tangent_prepare(fcfd_ctx* c) {
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    int rc = ensure_tangent(c);
    if (!rc) rc = ensure_adjoint(c);   // the load faces and the reduction buffers
    return rc;
}

// dR, dFX at fixed W in the direction of BC `index`'s values
static int param_tangent(fcfd_ctx* c, double* W_dev, int index, const fcfd_bc& d, double* dR_dev, hipStream_t s) {
    Params P = c->P;
    P.seed = make_seed(index, d);
    HIPCHK(c, hipMemsetAsync(c->jv_in, 0, sizeof(double) * (size_t)c->dg.nv * c->dg.ld, s));
    launch_param_tangent(c->dg, P, W_dev, c->jv_in, c->G, c->dG, c->PHI, c->dPHI, c->dref, c->dFX, dR_dev, c->status, s);
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

int fcfd_bc_tangent(fcfd_ctx* c, double* W_dev, int index, const fcfd_bc* dbc, double* dR_dev, double* dloads_host, void* stream) {
    if (!c || !W_dev || !dbc || !dR_dev) return FCFD_EINVAL;
    if (index < 0 || index >= c->opt.n_bc) return fail(c, FCFD_EINVAL, "BC index out of range");
    int rc = param_tangent_prepare(c);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);   // the primal ghosts of W
    if ((rc = param_tangent(c, W_dev, index, *dbc, dR_dev, s))) return rc;
    if (dloads_host) {
        launch_loads_tangent(c->dg, c->ag, c->dFX, c->opt.moment_center[0], c->opt.moment_center[1], c->opt.moment_center[2],
                             c->loads_tmp, c->n_loadfaces, c->loads_out, s);
        HIPCHK(c, hipGetLastError());
        if (c->opt.n_bc > 0)
            HIPCHK(c, hipMemcpyAsync(dloads_host, c->loads_out, sizeof(double) * 6 * c->opt.n_bc, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
    }
    return FCFD_OK;
}

int fcfd_vjp_params(fcfd_ctx* c, double* W_dev, const double* Rbar_dev, const double* fx_weight_dev, double* pbar_host, void* stream) {
    if (!c || !W_dev || !Rbar_dev || !pbar_host) return FCFD_EINVAL;
    int rc = param_tangent_prepare(c);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);
    const DevGrid& g = c->dg;
    const size_t n5 = (size_t)g.nv * g.ld;
    if (!c->red_part && ((rc = dalloc(c, (size_t)512 * 8, &c->red_part)) || (rc = dalloc(c, 64, &c->red_out)))) return rc;
    // the 11 values of each BC, one unit direction each
    for (int b = 0; b < c->opt.n_bc; ++b)
        for (int q = 0; q < FCFD_BC_NPARAMS; ++q) {
            fcfd_bc d{};
            double* slot[FCFD_BC_NPARAMS] = {&d.p, &d.T, &d.rho, &d.V[0], &d.V[1], &d.V[2], &d.p_total, &d.T_total, &d.mdot, &d.qdot, &d.nut};
            *slot[q] = 1.0;
            if ((rc = param_tangent(c, W_dev, b, d, c->jv_out, s))) return rc;
            double dot = 0.0, bf = 0.0;
            HIPCHK(c, dev_mdot(g, 1, c->jv_out, n5, Rbar_dev, c->red_part, c->red_out, &dot, s));
            if (fx_weight_dev) {
                launch_bface_dot(g, fx_weight_dev, c->dFX, c->red_part, c->red_out, s);
                HIPCHK(c, hipGetLastError());
                HIPCHK(c, hipMemcpyAsync(&bf, c->red_out, sizeof(double), hipMemcpyDeviceToHost, s));
                HIPCHK(c, hipStreamSynchronize(s));
            }
            pbar_host[(size_t)FCFD_BC_NPARAMS * b + q] = dot + bf;
        }
    return FCFD_OK;
}

int fcfd_jvp(fcfd_ctx* c, double* W_dev, double* dW_dev, double* dR_dev, int mode, void* stream) {
    if (!c || !W_dev || !dW_dev || !dR_dev) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    if (mode != FCFD_JAC_JVP_FULL && mode != FCFD_JAC_JVP_FROZEN) return fail(c, FCFD_EINVAL, "mode must be FCFD_JAC_JVP_FULL or _FROZEN");
    int rc = ensure_tangent(c);
    if (rc) return rc;
    launch_jvp(c->dg, c->P, W_dev, dW_dev, c->G, c->dG, c->PHI, c->dPHI, c->dref, c->dFX, dR_dev, mode == FCFD_JAC_JVP_FULL,
               c->status, (hipStream_t)stream, false, c->dSRC);
    HIPCHK(c, hipGetLastError());
    return FCFD_OK;
}

// buffers of the implicit step: residual, dt, updates, FD blocks, the
// Krylov basis of the requested restart and precision
static int ensure_newton(fcfd_ctx* c, const fcfd_newton_opts* o) {
    const DevGrid& g = c->dg;
    const size_t n5 = (size_t)g.nv * g.ld;
    const size_t b2 = 2 * (size_t)pair_groups(g.nv * g.nv);   // block entries in row pairs
    int rc;
    if (!c->R) {
        if ((rc = dalloc(c, n5, &c->R)) || (rc = dalloc(c, (size_t)g.ld, &c->dt)) || (rc = dalloc(c, n5, &c->dW)) ||
            (rc = dalloc(c, n5, &c->upd)) || (rc = dalloc(c, (size_t)512 * 8, &c->red_part)) || (rc = dalloc(c, 64, &c->red_out)))
            return rc;
    }
    if (!c->JL) {
        if ((rc = dalloc(c, b2 * g.F, &c->JL)) || (rc = dalloc(c, b2 * g.F, &c->JR)) ||
            (rc = dalloc(c, b2 * g.ld, &c->Dg)) || (rc = dalloc(c, n5, &c->zero_upd)))
            return rc;
    }
    const int fp32 = o->krylov_fp32 ? 1 : 0;
    const size_t m1 = (size_t)o->restart + 1;
    if (fp32 && c->gm_restart32 != o->restart) {   // (re)allocate the FP32 Krylov basis
        for (void* p : {(void*)c->w32.V, (void*)c->w32.z, (void*)c->w32.r, (void*)c->w32.part, (void*)c->w32.out_dev,
                        (void*)c->w32.coef_dev})
            dfree(c, p);
        if ((rc = dalloc(c, m1 * n5, &c->w32.V)) || (rc = dalloc(c, n5, &c->w32.z)) || (rc = dalloc(c, n5, &c->w32.r)) ||
            (rc = dalloc(c, m1 * 512, &c->w32.part)) || (rc = dalloc(c, m1, &c->w32.out_dev)) || (rc = dalloc(c, m1, &c->w32.coef_dev)))
            return rc;
        c->gm_restart32 = o->restart;
    }
    if (fp32 && !c->Dinv32) {
        if ((rc = dalloc(c, b2 * g.ld, &c->Dinv32)) || (rc = dalloc(c, b2 * g.ld, &c->D32)) ||
            (rc = dalloc(c, b2 * g.F, &c->JL32)) || (rc = dalloc(c, b2 * g.F, &c->JR32)) ||
            (rc = dalloc(c, n5, &c->b32)) || (rc = dalloc(c, n5, &c->x32)))
            return rc;
    }
    if (!fp32 && c->gm_restart64 != o->restart) {   // (re)allocate the FP64 Krylov basis
        for (void* p : {(void*)c->w64.V, (void*)c->w64.z, (void*)c->w64.r, (void*)c->w64.part, (void*)c->w64.out_dev,
                        (void*)c->w64.coef_dev})
            dfree(c, p);
        if ((rc = dalloc(c, m1 * n5, &c->w64.V)) || (rc = dalloc(c, n5, &c->w64.z)) || (rc = dalloc(c, n5, &c->w64.r)) ||
            (rc = dalloc(c, m1 * 512, &c->w64.part)) || (rc = dalloc(c, m1, &c->w64.out_dev)) || (rc = dalloc(c, m1, &c->w64.coef_dev)))
            return rc;
        c->gm_restart64 = o->restart;
    }
    if (!fp32 && !c->Dinv64 && (rc = dalloc(c, b2 * g.ld, &c->Dinv64))) return rc;
    return FCFD_OK;
}

// aggregation multigrid hierarchy on the interior face graph (host, once)
static int ensure_mg(fcfd_ctx* c) {
    if (c->mg.ready) return FCFD_OK;
    if (c->mg_tried) return fail(c, FCFD_ENOTSUP, "the mesh is too small for a multigrid hierarchy");
    c->mg_tried = true;
    const HostGrid& g = c->hg;
    std::vector<int> par(g.F), nei(g.F);
    for (int i = 0; i < g.F; ++i) {
        int f = g.ifperm[i];
        par[i] = g.perm[g.fparent[f]];
        nei[i] = g.perm[g.fneigh[f]];
    }
    int rc = mg_build(c->dg, par, nei, c->mg, 12, 256);
    if (rc) {
        mg_free(c->mg);
        return fail(c, rc, "multigrid set-up failed (HIP error)");
    }
    if (!c->mg.ready) return fail(c, FCFD_ENOTSUP, "the mesh is too small for a multigrid hierarchy");
    return FCFD_OK;
}

static bool graphs_enabled() {
    const char* e = std::getenv("FCFD_NO_GRAPHS");
    return !(e && std::strcmp(e, "1") == 0);
}

// capture one V-cycle (mg.fine_in -> mg.fine_out) with `sweeps` into a graph
static int ensure_mg_graph(fcfd_ctx* c, int sweeps) {
    if (c->mg_exec && c->mg_exec_sweeps == sweeps) return FCFD_OK;
    if (c->mg_exec) {
        HIPCHK(c, hipGraphExecDestroy(c->mg_exec));
        c->mg_exec = nullptr;
    }
    if (!c->cap_stream) HIPCHK(c, hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
    HIPCHK(c, hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeRelaxed));
    mg_vcycle(c->dg, c->colors, c->mg, c->Dg, c->JL, c->JR, c->Dinv64, c->mg.fine_in, c->mg.fine_out, sweeps, c->cap_stream);
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamEndCapture(c->cap_stream, &graph));
    hipError_t e = hipGraphInstantiate(&c->mg_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(c, e);
    c->mg_exec_sweeps = sweeps;
    return FCFD_OK;
}

// colouring of the interior cells on the face graph (color_graph: greedy in
// internal or breadth-first order, whichever needs fewer colours), for the
// multicolour block Gauss-Seidel sweeps
static int ensure_colors(fcfd_ctx* c) {
    if (c->colors.n > 0) return FCFD_OK;
    const HostGrid& g = c->hg;
    const int N = g.N;
    std::vector<int> hp, hf;
    half_faces(g, hp, hf);
    std::vector<int> par(g.F), nei(g.F);
    for (int i = 0; i < g.F; ++i) {
        int f = g.ifperm[i];
        par[i] = g.perm[g.fparent[f]];
        nei[i] = g.perm[g.fneigh[f]];
    }
    // interior face graph (ghost neighbours excluded), coloured greedily
    std::vector<int> aptr(N + 1, 0), adj;
    for (int i = 0; i < N; ++i) {
        for (int h = hp[i]; h < hp[i + 1]; ++h) {
            int fs = hf[h];
            int f = fs >= 0 ? fs : ~fs;
            int o = fs >= 0 ? nei[f] : par[f];
            if (o < N) adj.push_back(o);
        }
        aptr[i + 1] = (int)adj.size();
    }
    std::vector<int> ptr, cells;
    int ncol = 0;
    color_graph(N, aptr, adj, ptr, cells, ncol);
    int rc = upload(c, cells, &c->colors.cells);
    if (rc) return rc;
    c->colors.ptr = ptr;
    c->colors.n = ncol;
    return FCFD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// segregated two-equation RANS
// ---------------------------------------------------------------------------
namespace {

int kw_check(fcfd_ctx* c) {
    if (!c) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    if (!c->P.kw) return fail(c, FCFD_EINVAL, "the context has no two-equation RANS model (turbulence != FCFD_TURB_KW)");
    return FCFD_OK;
}

int kw_opts_check(fcfd_ctx* c, const fcfd_rans_opts* o) {
    if (o->restart < 1 || o->restart > 500 || o->max_iters < 0) return fail(c, FCFD_EINVAL, "bad RANS Krylov options");
    return FCFD_OK;
}

// the nv = 2 view of the grid for the Krylov kernels
DevGrid kw_grid(const fcfd_ctx* c) {
    DevGrid g2 = c->dg;
    g2.nv = 2;
    return g2;
}

int kw_ensure_krylov(fcfd_ctx* c, int restart) {
    if (c->kw_restart == restart) return FCFD_OK;
    for (void* p : {(void*)c->kw_w.V, (void*)c->kw_w.z, (void*)c->kw_w.r, (void*)c->kw_w.part, (void*)c->kw_w.out_dev,
                    (void*)c->kw_w.coef_dev})
        dfree(c, p);
    const size_t n2 = 2 * (size_t)c->dg.ld, m1 = (size_t)restart + 1;
    int rc;
    if ((rc = dalloc(c, m1 * n2, &c->kw_w.V)) || (rc = dalloc(c, n2, &c->kw_w.z)) || (rc = dalloc(c, n2, &c->kw_w.r)) ||
        (rc = dalloc(c, m1 * 512, &c->kw_w.part)) || (rc = dalloc(c, m1, &c->kw_w.out_dev)) || (rc = dalloc(c, m1, &c->kw_w.coef_dev)))
        return rc;
    c->kw_restart = restart;
    return FCFD_OK;
}

// NS ghosts, gradients and limiters of W (the RANS face loop reads them)
void kw_ns_prepare(fcfd_ctx* c, double* W, hipStream_t s) {
    launch_ghosts(c->dg, c->P, W, c->status, s);
    launch_grad_lim(c->dg, c->P, W, c->G, c->PHI, c->dref, s);
}

// RANS::solve at the NS state W (ghosts, gradients and limiters current)
int kw_solve(fcfd_ctx* c, double* W, const double* dt, const fcfd_rans_opts* o, fcfd_rans_stats* st, hipStream_t s) {
    int rc;
    if ((rc = kw_opts_check(c, o)) || (rc = kw_ensure_krylov(c, o->restart))) return rc;
    const DevGrid g2 = kw_grid(c);
    kw_launch_terms(c->dg, c->P, W, c->G, c->PHI, dt, c->kw, c->kwt, c->hg.fperm[c->hg.F - 1], c->status, s);
    dev_bj_invert<double>(g2, c->kwt.D, c->kw_Dinv, c->status, s);
    LinOp<double> A{[&](const double* x, double* y, hipStream_t ss) { dev_bsr_matvec<double>(g2, c->kwt.D, c->kwt.JL, c->kwt.JR, x, y, ss); }};
    LinOp<double> M{[&](const double* x, double* y, hipStream_t ss) { dev_bj_apply<double>(g2, c->kw_Dinv, x, y, ss); }};
    GmresOpts go;
    go.restart = o->restart; go.max_iters = o->max_iters; go.rtol = o->rtol; go.abstol = o->abstol;
    go.knoll = 1;   // KSPSetInitialGuessKnoll (rans_petsc_functions.cc:76)
    GmresStats gs;
    rc = gmres<double>(g2, A, M, c->kwt.rhs, c->kw_dU, go, c->kw_w, gs, s);
    if (rc) return fail(c, rc, "RANS GMRES failed (HIP error)");
    double out[3];
    HIPCHK(c, kw_update(c->dg, c->P, c->kw, c->kw_dU, dt, c->kw_part, c->kw_out, out, s));
    if (st) {
        st->krylov_iters = gs.iters; st->krylov_converged = gs.converged; st->krylov_rnorm = gs.rnorm; st->krylov_r0norm = gs.r0norm;
        st->norms[0] = std::sqrt(out[0]); st->norms[1] = std::sqrt(out[1]);
    }
    if (out[2] > 0.0) return fail(c, FCFD_EDIVERGED, "Divergence detected in RANS (rans.cc:195-200)");
    kw_launch_refresh(c->dg, c->P, W, c->kw, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(s));
    int bad = 0;
    HIPCHK(c, hipMemcpy(&bad, c->status, sizeof(int), hipMemcpyDeviceToHost));
    if (bad & 2) return fail(c, FCFD_EDIVERGED, "singular diagonal block in the RANS block-Jacobi preconditioner");
    return FCFD_OK;
}

}  // namespace

extern "C" {

int fcfd_rans_init(fcfd_ctx* c, double* W_dev, double intensity, double visc_ratio, const fcfd_rans_opts* opts, void* stream) {
    int rc = kw_check(c);
    if (rc) return rc;
    if (!W_dev) return FCFD_EINVAL;
    if (opts && (rc = kw_opts_check(c, opts))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int nbc = c->opt.n_bc;
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);
    kw_launch_init(c->dg, c->P, W_dev, intensity, visc_ratio, c->kw, s);
    kw_launch_inlets(c->dg, c->P, nbc, W_dev, c->kw_f0, c->kw_inl, s);
    std::vector<double> h(3 * (size_t)std::max(nbc, 1), 0.0);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(h.data(), c->kw_inl, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->kw_bcv = h;
    set_params(c);
    // the BC table is read by every kernel: replace it with no work in flight
    HIPCHK(c, hipDeviceSynchronize());
    HIPCHK(c, hipMemcpy((void*)c->dg.bc, c->bch.data(), sizeof(BcDev) * c->bch.size(), hipMemcpyHostToDevice));
    kw_launch_refresh(c->dg, c->P, W_dev, c->kw, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(s));
    if (opts) c->kw_opts = *opts;
    c->kw_ready = true;
    return FCFD_OK;
}

int fcfd_rans_set(fcfd_ctx* c, double* W_dev, const double* K_host, void* stream) {
    int rc = kw_check(c);
    if (rc) return rc;
    if (!W_dev || !K_host) return FCFD_EINVAL;
    if (c->kw_bcv.empty()) return fail(c, FCFD_EINVAL, "fcfd_rans_set needs fcfd_rans_init first (the inlet values)");
    const HostGrid& hg = c->hg;
    const size_t ld = c->dg.ld;
    std::vector<double> h(2 * ld, 0.0);
    HIPCHK(c, hipDeviceSynchronize());
    HIPCHK(c, hipMemcpy(h.data(), c->kw.K, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < hg.N; ++r) {
        const int i = hg.perm[r];
        h[i] = K_host[2 * (size_t)r];
        h[ld + i] = K_host[2 * (size_t)r + 1];
    }
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(c, hipMemcpyAsync(c->kw.K, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
    launch_ghosts(c->dg, c->P, W_dev, c->status, s);
    kw_launch_refresh(c->dg, c->P, W_dev, c->kw, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(s));
    c->kw_ready = true;
    return FCFD_OK;
}

int fcfd_rans_get(fcfd_ctx* c, double* K_host, double* mut_host, double* G_host, double* strain_host, double* mutf_host) {
    int rc = kw_check(c);
    if (rc) return rc;
    const HostGrid& hg = c->hg;
    const size_t ld = c->dg.ld;
    const int T = hg.N + hg.B;
    HIPCHK(c, hipDeviceSynchronize());
    std::vector<double> K(2 * ld), mut(ld), G(6 * ld), st(ld);
    HIPCHK(c, hipMemcpy(K.data(), c->kw.K, K.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(mut.data(), c->kw.mut, mut.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(G.data(), c->kw.GK, G.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(st.data(), c->kw.strain, st.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < T; ++r) {
        const int i = hg.perm[r];
        if (K_host) { K_host[2 * (size_t)r] = K[i]; K_host[2 * (size_t)r + 1] = K[ld + i]; }
        if (mut_host) mut_host[r] = mut[i];
        if (strain_host) strain_host[r] = st[i];
        if (G_host && r < hg.N)
            for (int q = 0; q < 6; ++q) G_host[6 * (size_t)r + q] = G[q * ld + i];
    }
    if (mutf_host) {   // mu_t.face: the BC's value where fixed, else the face average (Variable::face_calculate)
        for (int f = 0; f < hg.F; ++f) {
            const int b = hg.fbc[f];
            if (b >= 0 && c->bch[b].kw_mfix) { mutf_host[f] = c->bch[b].kw_mut; continue; }
            double t = 0.;
            for (int k = hg.aptr[f]; k < hg.aptr[f + 1]; ++k) t += hg.aw[k] * mut[hg.perm[hg.aidx[k]]];
            mutf_host[f] = t;
        }
    }
    return FCFD_OK;
}

int fcfd_rans_terms(fcfd_ctx* c, double* W_dev, const double* dt_dev, double* rhs_host, double* diag_host, double* jf_host,
                    void* stream) {
    int rc = kw_check(c);
    if (rc) return rc;
    if (!W_dev || !dt_dev) return FCFD_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    kw_ns_prepare(c, W_dev, s);
    kw_launch_terms(c->dg, c->P, W_dev, c->G, c->PHI, dt_dev, c->kw, c->kwt, c->hg.fperm[c->hg.F - 1], c->status, s);
    HIPCHK(c, hipGetLastError());
    const HostGrid& hg = c->hg;
    const size_t ld = c->dg.ld, F = hg.F;
    std::vector<double> rhs(2 * ld), D(4 * ld), FV(6 * F);
    HIPCHK(c, hipMemcpyAsync(rhs.data(), c->kwt.rhs, rhs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(D.data(), c->kwt.D, D.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(FV.data(), c->kwt.FV, FV.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    for (int r = 0; r < hg.N; ++r) {
        const int i = hg.perm[r];
        if (rhs_host) { rhs_host[2 * (size_t)r] = rhs[i]; rhs_host[2 * (size_t)r + 1] = rhs[ld + i]; }
        if (diag_host)
            for (int q = 0; q < 4; ++q) diag_host[4 * (size_t)r + q] = D[pidx(ld, q, i)];
    }
    if (jf_host)
        for (int f = 0; f < hg.F; ++f) {
            const int i = hg.fperm[f];
            for (int q = 0; q < 6; ++q) jf_host[6 * (size_t)f + q] = FV[q * F + i];
        }
    return FCFD_OK;
}

int fcfd_rans_step(fcfd_ctx* c, double* W_dev, const double* dt_dev, const fcfd_rans_opts* o, fcfd_rans_stats* st, void* stream) {
    int rc = kw_check(c);
    if (rc) return rc;
    if (!W_dev || !dt_dev || !o) return FCFD_EINVAL;
    if (!c->kw_ready) return fail(c, FCFD_EINVAL, "fcfd_rans_step needs fcfd_rans_init or fcfd_rans_set first");
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(c, hipMemsetAsync(c->status, 0, sizeof(int), s));
    kw_ns_prepare(c, W_dev, s);
    return kw_solve(c, W_dev, dt_dev, o, st, s);
}

}  // extern "C"

// the neighbour cells (ghosts included) of the interior cells in internal
// numbering, for the WS95 pressure-difference floor
static int ensure_neighbours(fcfd_ctx* c) {
    if (c->nb_ptr) return FCFD_OK;
    const HostGrid& g = c->hg;
    std::vector<int> p(g.N + 1, 0), idx;
    for (int i = 0; i < g.N; ++i) {
        const int r = g.iperm[i];
        for (int k = g.nbptr[r]; k < g.nbptr[r + 1]; ++k) idx.push_back(g.perm[g.nb[k]]);
        p[i + 1] = (int)idx.size();
    }
    int rc;
    if ((rc = upload(c, p, &c->nb_ptr)) || (rc = upload(c, idx, &c->nb_idx))) return rc;
    return FCFD_OK;
}

static int newton_impl(fcfd_ctx* c, double* W_dev, const fcfd_newton_opts* o, const fcfd_pseudo_opts* ps, fcfd_newton_stats* st,
                       void* stream) {
    if (!c || !W_dev || !o) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    int rc0 = 0;
    if ((rc0 = check_dt(c, o->dt_type, o->cfl))) return rc0;
    if (ps) {
        if (ps->dtau_type != FCFD_DT_FIXED && ps->dtau_type != FCFD_DT_CFL_MAX && ps->dtau_type != FCFD_DT_CFL_LOCAL)
            return fail(c, FCFD_EINVAL, "unknown pseudo time step type");
        if (!(ps->dtau > 0.0)) return fail(c, FCFD_EINVAL, "the pseudo step size or CFL number must be positive");
        if (ps->preconditioner != FCFD_PSEUDO_PC_NONE && ps->preconditioner != FCFD_PSEUDO_PC_WS95)
            return fail(c, FCFD_EINVAL, "unknown pseudo-time preconditioner");
        if (ps->ps_step < 1) return fail(c, FCFD_EINVAL, "pseudo steps count from 1");
        if (ps->ps_step > 1 && !ps->Wn_dev) return fail(c, FCFD_EINVAL, "a pseudo step after the first needs the state at time level n");
        if (c->P.kw && c->kw_ready)
            return fail(c, FCFD_ENOTSUP, "dual time stepping with the segregated two-equation RANS model is not supported");
        if (!c->dtau && ((rc0 = dalloc(c, (size_t)c->dg.ld, &c->dtau)) || (rc0 = dalloc(c, (size_t)c->dg.ld, &c->drdp)))) return rc0;
        if (ps->preconditioner == FCFD_PSEUDO_PC_WS95 && (rc0 = ensure_neighbours(c))) return rc0;
    }
    if (o->jacobian != FCFD_JAC_FD && o->jacobian != FCFD_JAC_JVP_FULL && o->jacobian != FCFD_JAC_JVP_FROZEN)
        return fail(c, FCFD_EINVAL, "unknown Jacobian mode");
    const bool mfree = o->jacobian != FCFD_JAC_FD;
    if (mfree && (rc0 = ensure_tangent(c))) return rc0;
    // FP32 Krylov + exact tangent: the tangent's derivative lane in FP32 too
    // (FP64 lane behind conversions where Green-Gauss cells need it, or with
    // FCFD_TANGENT_FP64=1, the comparison case of the tests)
    const char* t64 = std::getenv("FCFD_TANGENT_FP64");
    const bool tan32 = mfree && o->krylov_fp32 && !c->dg.ggc && !(t64 && t64[0] == '1');
    if (tan32 && (rc0 = ensure_tangent32(c))) return rc0;
    if (o->restart < 1 || o->restart > 500 || o->max_iters < 0) return fail(c, FCFD_EINVAL, "bad Krylov options");
    if (o->precond != FCFD_PC_BLOCK_JACOBI && o->precond != FCFD_PC_BLOCK_SGS && o->precond != FCFD_PC_MULTIGRID)
        return fail(c, FCFD_EINVAL, "unknown preconditioner");
    const bool mgp = o->precond == FCFD_PC_MULTIGRID;
    const bool sgs = o->precond == FCFD_PC_BLOCK_SGS;
    const int sweeps = std::max(o->sweeps, 1);
    if ((sgs || mgp) && (rc0 = ensure_colors(c))) return rc0;
    if (mgp && (rc0 = ensure_mg(c))) return rc0;
    const DevGrid& g = c->dg;
    hipStream_t s = (hipStream_t)stream;
    const size_t n5 = (size_t)g.nv * g.ld;      // one [nv][ld] vector
    const size_t b2 = 2 * (size_t)pair_groups(g.nv * g.nv);   // entries of one block in row pairs (padded)
    int rc = ensure_newton(c, o);
    if (rc) return rc;
    const int fp32 = o->krylov_fp32 ? 1 : 0;
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    fcfd_newton_stats S{};
    // the step reports the BC and pivot flags its own kernels raise
    HIPCHK(c, hipMemsetAsync(c->status, 0, sizeof(int), s));
    auto t0 = clk::now();
    // 1. residual and its norm
    launch_ghosts(g, c->P, W_dev, c->status, s);
    launch_grad_lim(g, c->P, W_dev, c->G, c->PHI, c->dref, s, c->SRC);
    launch_flux(g, c->P, W_dev, c->G, c->PHI, c->FX, c->R, c->status, s);
    launch_residual_sum(g, c->SRC, c->FX, c->R, s);
    launch_time_step(g, c->P, W_dev, o->dt_type, o->cfl, c->dt, c->dt_red, s);
    if (ps) {   // dual time: the pseudo step, its block and the right-hand side R - (P V/dt)(W - Wn)
        launch_time_step(g, c->P, W_dev, ps->dtau_type, ps->dtau, c->dtau, c->dt_red, s);
        launch_pseudo_rhs(g, c->P, W_dev, ps->Wn_dev, c->dt, c->nb_ptr, c->nb_idx, ps->preconditioner == FCFD_PSEUDO_PC_WS95,
                          ps->ps_step > 1, c->R, c->drdp, s);
    }
    double rr = 0.0;
    HIPCHK(c, dev_mdot(g, 1, c->R, n5, c->R, c->red_part, c->red_out, &rr, s));
    S.res_norm = std::sqrt(rr);
    auto t1 = clk::now();
    // 2. FD Jacobian + time terms (+ the pseudo-time block), block-Jacobi inverse
    launch_fd_jacobian(g, c->P, W_dev, c->G, c->PHI, c->upd, c->dt, c->JL, c->JR, c->Dg, c->status, s);
    if (ps) launch_pseudo_diag(g, c->P, W_dev, c->dtau, c->drdp, c->Dg, s);
    if (fp32) {
        dev_bj_invert<float>(g, c->Dg, c->Dinv32, c->status, s);
        if (!mfree || sgs) {
            dev_convert<double, float>(b2 * g.ld, c->Dg, c->D32, s);
            dev_convert<double, float>(b2 * g.F, c->JL, c->JL32, s);
            dev_convert<double, float>(b2 * g.F, c->JR, c->JR32, s);
        }
        dev_convert<double, float>(n5, c->R, c->b32, s);
    } else if (!mgp) {
        dev_bj_invert<double>(g, c->Dg, c->Dinv64, c->status, s);
    }
    // the assembled operator's matvec streams the off-diagonal blocks from
    // the sliced layout (FCFD_BSR_MATVEC=1: from the per-face blocks, the
    // comparison case of the tests)
    const char* bsr = std::getenv("FCFD_BSR_MATVEC");
    const bool sell = !mfree && !(bsr && bsr[0] == '1');
    if (sell) {
        if ((rc = ensure_sell(c, fp32))) return rc;
        if (fp32) dev_sell_pack<float>(g, c->sell, c->JL32, c->JR32, c->sell32, s);
        else dev_sell_pack<double>(g, c->sell, c->JL, c->JR, c->sell64, s);
    }
    const bool use_graph = mgp && graphs_enabled();
    if (mgp) {   // FP64 V-cycle on the FD blocks: fine inverse + coarse levels
        if (!c->Dinv64 && (rc = dalloc(c, b2 * g.ld, &c->Dinv64))) return rc;
        mg_setup(g, c->mg, c->Dg, c->JL, c->JR, c->Dinv64, c->status, s);
        if (use_graph && (rc = ensure_mg_graph(c, sweeps))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(s));
    auto t2 = clk::now();
    // 3. GMRES
    GmresOpts go;
    go.restart = o->restart; go.max_iters = o->max_iters; go.rtol = o->rtol; go.abstol = o->abstol; go.knoll = o->knoll_guess;
    GmresStats gs;
    // matrix-free operator: y = (P V/dt) x - J x with the tangent of W (the
    // ghost slots of W are rewritten with the same values by each tangent
    // call).  The primal ghosts, gradients and limiters at W are those of
    // step 1, so the tangent reads them (FCFD_JVP_RECOMPUTE=1: recomputes
    // them in every call, the comparison case of the tests).
    const char* jrc = std::getenv("FCFD_JVP_RECOMPUTE");
    const bool known = !(jrc && jrc[0] == '1');
    // the FP32 operator's limiter linearisation at this W (read by every product)
    if (tan32 && known && o->jacobian == FCFD_JAC_JVP_FULL && c->lin32.nb) launch_lim_lin(g, c->P, W_dev, c->G, c->dref, c->lin32, s);
    auto jvp_op = [&](const double* x, double* y, hipStream_t ss) {
        (void)hipMemcpyAsync(c->jv_in, x, n5 * sizeof(double), hipMemcpyDeviceToDevice, ss);
        launch_jvp(g, c->P, W_dev, c->jv_in, c->G, c->dG, c->PHI, c->dPHI, c->dref, c->dFX, c->jv_R,
                   o->jacobian == FCFD_JAC_JVP_FULL, c->status, ss, known, c->dSRC);
        launch_newton_op(g, c->P, W_dev, c->dt, c->jv_in, c->jv_R, y, ss);   // jv_in: x may alias y
        if (ps) launch_pseudo_op(g, c->P, W_dev, c->dtau, c->drdp, c->jv_in, y, ss);
    };
    // preconditioners: block Jacobi, block SGS, or the FP64 multigrid V-cycle
    // (FP32 Krylov vectors converted at its boundary)
    auto prec64 = [&](const double* x, double* y, hipStream_t ss) {
        if (use_graph) {
            (void)hipMemcpyAsync(c->mg.fine_in, x, n5 * sizeof(double), hipMemcpyDeviceToDevice, ss);
            (void)hipGraphLaunch(c->mg_exec, ss);
            (void)hipMemcpyAsync(y, c->mg.fine_out, n5 * sizeof(double), hipMemcpyDeviceToDevice, ss);
        } else if (mgp) mg_vcycle(g, c->colors, c->mg, c->Dg, c->JL, c->JR, c->Dinv64, x, y, sweeps, ss);
        else if (sgs) dev_bsgs_apply<double>(g, c->colors, c->Dinv64, c->JL, c->JR, x, y, sweeps, ss);
        else dev_bj_apply<double>(g, c->Dinv64, x, y, ss);
    };
    auto prec32 = [&](const float* x, float* y, hipStream_t ss) {
        if (mgp) {
            dev_convert<float, double>(n5, x, c->mg.fine_in, ss);
            if (use_graph) (void)hipGraphLaunch(c->mg_exec, ss);
            else mg_vcycle(g, c->colors, c->mg, c->Dg, c->JL, c->JR, c->Dinv64, c->mg.fine_in, c->mg.fine_out, sweeps, ss);
            dev_convert<double, float>(n5, c->mg.fine_out, y, ss);
        } else if (sgs) {
            dev_bsgs_apply<float>(g, c->colors, c->Dinv32, c->JL32, c->JR32, x, y, sweeps, ss);
        } else {
            dev_bj_apply<float>(g, c->Dinv32, x, y, ss);
        }
    };
    // the FP32 operator: y = (P V/dt) x - J x (+ the pseudo-time term) with the
    // tangent's derivative lane in FP32 (values FP64, from W)
    auto jvp_op32 = [&](const float* x, float* y, hipStream_t ss) {
        (void)hipMemcpyAsync(c->jv_in32, x, n5 * sizeof(float), hipMemcpyDeviceToDevice, ss);
        launch_jvp(g, c->P, W_dev, c->jv_in32, c->G, c->dG32, c->PHI, c->dPHI32, c->dref, c->dFX32, c->jv_R32,
                   o->jacobian == FCFD_JAC_JVP_FULL, c->status, ss, known, c->dSRC32, c->lin32.nb ? &c->lin32 : nullptr);
        launch_newton_op(g, c->P, W_dev, c->dt, c->jv_in32, c->jv_R32, y, ss);
        if (ps) launch_pseudo_op(g, c->P, W_dev, c->dtau, c->drdp, c->jv_in32, y, ss);
    };
    if (fp32 && mfree) {
        LinOp<float> A{[&](const float* x, float* y, hipStream_t ss) {
            if (tan32) {
                jvp_op32(x, y, ss);
            } else {
                dev_convert<float, double>(n5, x, c->jv_out, ss);
                jvp_op(c->jv_out, c->jv_out, ss);
                dev_convert<double, float>(n5, c->jv_out, y, ss);
            }
        }};
        LinOp<float> M{prec32};
        rc = gmres<float>(g, A, M, c->b32, c->x32, go, c->w32, gs, s);
        if (!rc) dev_convert<float, double>(n5, c->x32, c->dW, s);
    } else if (mfree) {
        LinOp<double> A{[&](const double* x, double* y, hipStream_t ss) { jvp_op(x, y, ss); }};
        LinOp<double> M{prec64};
        rc = gmres<double>(g, A, M, c->R, c->dW, go, c->w64, gs, s);
    } else if (fp32) {
        LinOp<float> A{[&](const float* x, float* y, hipStream_t ss) {
            if (sell) dev_sell_matvec<float>(g, c->sell, c->D32, c->sell32, x, y, ss);
            else dev_bsr_matvec<float>(g, c->D32, c->JL32, c->JR32, x, y, ss);
        }};
        LinOp<float> M{prec32};
        rc = gmres<float>(g, A, M, c->b32, c->x32, go, c->w32, gs, s);
        if (!rc) dev_convert<float, double>(n5, c->x32, c->dW, s);
    } else {
        LinOp<double> A{[&](const double* x, double* y, hipStream_t ss) {
            if (sell) dev_sell_matvec<double>(g, c->sell, c->Dg, c->sell64, x, y, ss);
            else dev_bsr_matvec<double>(g, c->Dg, c->JL, c->JR, x, y, ss);
        }};
        LinOp<double> M{prec64};
        rc = gmres<double>(g, A, M, c->R, c->dW, go, c->w64, gs, s);
    }
    if (rc) return fail(c, rc, "GMRES failed (HIP error)");
    HIPCHK(c, hipStreamSynchronize(s));
    // 3b. the segregated two-equation model's solve at the same NS state
    //     (NavierStokes::solve calls RANS::solve after petsc_solve and before
    //     update_variables, ns.cc:106-110)
    if (c->P.kw && c->kw_ready) {
        fcfd_rans_stats rs{};
        rc = kw_solve(c, W_dev, c->dt, &c->kw_opts, &rs, s);
        S.rans_iters = rs.krylov_iters; S.rans_norms[0] = rs.norms[0]; S.rans_norms[1] = rs.norms[1];
        if (rc) {
            if (st) *st = S;
            return rc;
        }
    }
    auto t3 = clk::now();
    // 4. update W (relaxed if it would change p or T by more than
    //    max_rel_update), norms, NaN check; keep dW as the previous update
    S.relax = 1.0;
    if (o->max_rel_update > 0.0) {
        double rmax = 0.0;
        HIPCHK(c, dev_relmax(g, c->P.Pref, c->P.Tref, W_dev, c->dW, c->red_part, &rmax, s));
        if (rmax > o->max_rel_update) {
            S.relax = o->max_rel_update / rmax;
            dev_scale_copy<double>(n5, S.relax, c->dW, c->dW, s);
        }
    }
    double out[5];
    HIPCHK(c, dev_update(g, W_dev, c->dW, c->dt, c->red_part, c->red_out, out, s));
    if (!(out[3] > 0.0)) HIPCHK(c, hipMemcpyAsync(c->upd, c->dW, n5 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipStreamSynchronize(s));
    auto t4 = clk::now();
    S.update_norms[0] = std::sqrt(out[0]); S.update_norms[1] = std::sqrt(out[1]); S.update_norms[2] = std::sqrt(out[2]);
    S.nut_norm = std::sqrt(out[4]);
    S.krylov_iters = gs.iters; S.krylov_converged = gs.converged; S.krylov_rnorm = gs.rnorm; S.krylov_r0norm = gs.r0norm;
    S.ms_residual = ms(t0, t1); S.ms_jacobian = ms(t1, t2); S.ms_krylov = ms(t2, t3); S.ms_update = ms(t3, t4);
    if (st) *st = S;
    if (out[3] > 0.0) return fail(c, FCFD_EDIVERGED, "Divergence detected: NaN/Inf in the update (ns.cc:350-355)");
    int bad = 0;
    HIPCHK(c, hipMemcpy(&bad, c->status, sizeof(int), hipMemcpyDeviceToHost));
    if (bad & 2) return fail(c, FCFD_EDIVERGED, "singular diagonal block in the block-Jacobi preconditioner");
    if (bad & 1)   // the reference exits here (ns_apply_bcs.cc:54-59)
        return fail(c, FCFD_EBC, "an inlet became supersonic; the thermodynamic state must be specified (ns_apply_bcs.cc:54-59)");
    return FCFD_OK;
}

extern "C" {

int fcfd_newton_step(fcfd_ctx* c, double* W_dev, const fcfd_newton_opts* o, fcfd_newton_stats* st, void* stream) {
    return newton_impl(c, W_dev, o, nullptr, st, stream);
}

int fcfd_pseudo_step(fcfd_ctx* c, double* W_dev, const fcfd_newton_opts* o, const fcfd_pseudo_opts* ps, fcfd_newton_stats* st,
                     void* stream) {
    if (!ps) return fail(c, FCFD_EINVAL, "null pseudo-time options");
    return newton_impl(c, W_dev, o, ps, st, stream);
}

int fcfd_adjoint_step(fcfd_ctx* c, double* W_dev, const double* g_dev, double* lam_dev, const fcfd_newton_opts* o,
                      fcfd_newton_stats* st, void* stream) {
    if (!c || !W_dev || !g_dev || !lam_dev || !o) return FCFD_EINVAL;
    if (!c->device) return fail(c, FCFD_EINVAL, "context has no device data");
    if (int rc_dt = check_dt(c, o->dt_type, o->cfl)) return rc_dt;
    if (o->krylov_fp32) return fail(c, FCFD_ENOTSUP, "the adjoint step solves in FP64 (krylov_fp32 must be 0)");
    if (o->precond != FCFD_PC_BLOCK_JACOBI && o->precond != FCFD_PC_BLOCK_SGS && o->precond != FCFD_PC_MULTIGRID)
        return fail(c, FCFD_EINVAL, "unknown preconditioner");
    if (o->precond == FCFD_PC_BLOCK_SGS && o->sweeps < 1) return fail(c, FCFD_EINVAL, "block SGS needs sweeps >= 1");
    if (o->restart < 1 || o->restart > 500 || o->max_iters < 0) return fail(c, FCFD_EINVAL, "bad Krylov options");
    const bool sgs = o->precond == FCFD_PC_BLOCK_SGS, mgp = o->precond == FCFD_PC_MULTIGRID;
    int rc = ensure_adjoint(c);
    if (!rc) rc = ensure_newton(c, o);
    if (!rc && (sgs || mgp)) rc = ensure_colors(c);
    if (!rc && mgp) rc = ensure_mg(c);
    if (rc) return rc;
    const DevGrid& g = c->dg;
    hipStream_t s = (hipStream_t)stream;
    const size_t n5 = (size_t)g.nv * g.ld;
    if (!c->adj_v && (rc = dalloc(c, n5, &c->adj_v))) return rc;
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    fcfd_newton_stats S{};
    // the step reports the pivot flags its own kernels raise
    HIPCHK(c, hipMemsetAsync(c->status, 0, sizeof(int), s));
    auto t0 = clk::now();
    // 1. adjoint residual rho = J^T lam + g (the VJP also refreshes ghosts, G and PHI of W)
    launch_vjp(g, c->ag, c->P, W_dev, c->G, c->PHI, c->dref, lam_dev, nullptr, c->aw, c->R, c->status, s);
    dev_axpy<double>(n5, 1.0, g_dev, c->R, s);
    double rr = 0.0;
    HIPCHK(c, dev_mdot(g, 1, c->R, n5, c->R, c->red_part, c->red_out, &rr, s));
    S.res_norm = std::sqrt(rr);
    auto t1 = clk::now();
    if (!std::isfinite(S.res_norm)) {
        if (st) *st = S;
        return fail(c, FCFD_EDIVERGED, "Divergence detected: NaN/Inf in the adjoint residual");
    }
    // 2. local time step, FD blocks with the floor step, their diagonal
    //    inverses (transposed block Jacobi, transposed block SGS, or the
    //    multigrid hierarchy's transposed V-cycle)
    launch_time_step(g, c->P, W_dev, o->dt_type, o->cfl, c->dt, c->dt_red, s);
    launch_fd_jacobian(g, c->P, W_dev, c->G, c->PHI, c->zero_upd, c->dt, c->JL, c->JR, c->Dg, c->status, s);
    dev_bj_invert<double>(g, c->Dg, c->Dinv64, c->status, s);
    if (mgp) mg_setup(g, c->mg, c->Dg, c->JL, c->JR, c->Dinv64, c->status, s);   // the hierarchy's blocks (used transposed)
    HIPCHK(c, hipStreamSynchronize(s));
    auto t2 = clk::now();
    // 3. GMRES on ((P V/dt) - J)^T delta = rho
    GmresOpts go;
    go.restart = o->restart; go.max_iters = o->max_iters; go.rtol = o->rtol; go.abstol = o->abstol; go.knoll = o->knoll_guess;
    GmresStats gs;
    LinOp<double> A{[&](const double* x, double* y, hipStream_t ss) {
        launch_vjp(g, c->ag, c->P, W_dev, c->G, c->PHI, c->dref, x, nullptr, c->aw, c->adj_v, c->status, ss);
        launch_newton_op_t(g, c->P, W_dev, c->dt, x, c->adj_v, y, ss);
    }};
    LinOp<double> M{[&](const double* x, double* y, hipStream_t ss) {
        if (mgp) mg_vcycle(g, c->colors, c->mg, c->Dg, c->JL, c->JR, c->Dinv64, x, y, std::max(o->sweeps, 1), ss, true);
        else if (sgs) dev_bsgs_apply<double>(g, c->colors, c->Dinv64, c->JL, c->JR, x, y, o->sweeps, ss, true, 0, true);
        else dev_bj_apply_t(g, c->Dinv64, x, y, ss);
    }};
    rc = gmres<double>(g, A, M, c->R, c->dW, go, c->w64, gs, s);
    if (rc) return fail(c, rc, "GMRES failed (HIP error)");
    HIPCHK(c, hipStreamSynchronize(s));
    auto t3 = clk::now();
    // 4. lam += delta
    dev_axpy<double>(n5, 1.0, c->dW, lam_dev, s);
    S.relax = 1.0;
    HIPCHK(c, hipStreamSynchronize(s));
    auto t4 = clk::now();
    S.krylov_iters = gs.iters; S.krylov_converged = gs.converged; S.krylov_rnorm = gs.rnorm; S.krylov_r0norm = gs.r0norm;
    S.ms_residual = ms(t0, t1); S.ms_jacobian = ms(t1, t2); S.ms_krylov = ms(t2, t3); S.ms_update = ms(t3, t4);
    if (st) *st = S;
    int bad = 0;
    HIPCHK(c, hipMemcpy(&bad, c->status, sizeof(int), hipMemcpyDeviceToHost));
    if (bad & 2) return fail(c, FCFD_EDIVERGED, "singular diagonal block in the block-Jacobi preconditioner");
    return FCFD_OK;
}

int fcfd_set_bc(fcfd_ctx* c, int index, const fcfd_bc* bc) {
    if (!c || !bc) return FCFD_EINVAL;
    if (index < 0 || index >= c->opt.n_bc) return fail(c, FCFD_EINVAL, "BC index out of range");
    const fcfd_bc& old = c->opt.bc[index];
    if (bc->type != old.type || bc->kind != old.kind || bc->specified != old.specified || bc->thermal != old.thermal)
        return fail(c, FCFD_EBC, "fcfd_set_bc changes BC values only: type, kind, specified and thermal are fixed at create");
    const fcfd_options saved = c->opt;
    c->opt.bc[index] = *bc;
    int rc = validate_options(c);
    if (rc) {
        c->opt = saved;
        return rc;
    }
    const int viscous_before = c->P.viscous;
    set_params(c);
    if (c->P.viscous != viscous_before) {   // the kernels and adjoint workspace were chosen for the old flag
        c->opt = saved;
        set_params(c);
        return fail(c, FCFD_EBC, "fcfd_set_bc cannot switch the diffusive flux on or off (a FIXED_Q wall's qdot between 0 and non-zero on an inviscid case)");
    }
    if (c->device) {
        // the table is read by every kernel: replace it with no work in flight
        HIPCHK(c, hipDeviceSynchronize());
        HIPCHK(c, hipMemcpy((void*)c->dg.bc, c->bch.data(), sizeof(BcDev) * c->bch.size(), hipMemcpyHostToDevice));
    }
    return FCFD_OK;
}

}  // extern "C"
