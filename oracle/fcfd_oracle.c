/*
 * fcfd_oracle.c -- TEST INFRASTRUCTURE ONLY (the parity checker).
 *
 * A line-by-line CPU restatement, in plain C and FP64, of Free CFD 2.0's
 * unstructured finite-volume residual and FD-Jacobian path.  Every function
 * names the reference file:line it follows (paths relative to the reference
 * repository root).  Built with gcc -O2 -ffp-contract=off so that the
 * arithmetic sequence is the reference's (x86-64 GCC -O3 emits no FMA).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this code; the product library never links it.
 *
 * Decisions where the reference is undefined:
 *  D1  update_boundaries() has no source (declared ns.h:165).  Ghost values
 *      are set per boundary face by apply_bcs() with the parent cell-centre
 *      state as `left` and `right` pre-set to a copy of `left`; the ghost
 *      receives right.{p,V,T}_center.
 *  D2  outlet() "extrapolate everything" branch (ns_apply_bcs.cc:189-194)
 *      leaves right.rho at a stale value from the previous face; here
 *      right.rho = EOS(right.p,right.T) (== left.rho).
 *  D3  curvilinear_grad_map() erases std::set elements it is iterating
 *      (curvilinear_grad_map.cc:57-78); here the intended greedy pairing is
 *      used: take the lowest remaining face slot, pair it with the most
 *      opposite remaining later slot (first minimum wins), remove both.
 *  D4  (none: repeated-node cells follow the reference, including its
 *      hexahedron -> prism node rebuild, grid_create_elements.cc:426-467)
 *  D5  (none: WTLI fits of any stencil size are restated; the device path
 *      refuses stencils wider than two cells)
 */
#include "fcfd_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Vec3D (src/vec3d/vec3d.cc:29-211): same operation order            */
/* ------------------------------------------------------------------ */
typedef struct { double c[3]; } V3;
static V3 v3(double x, double y, double z) { V3 r; r.c[0] = x; r.c[1] = y; r.c[2] = z; return r; }
static V3 vadd(V3 a, V3 b) { return v3(a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]); }
static V3 vsub(V3 a, V3 b) { return v3(a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]); }
static V3 vsc(double s, V3 a) { return v3(s * a.c[0], s * a.c[1], s * a.c[2]); }     /* operator*(double,Vec3D) */
static V3 vmul(V3 a, double s) { return v3(a.c[0] * s, a.c[1] * s, a.c[2] * s); }     /* Vec3D*double */
static V3 vdiv(V3 a, double s) { return v3(a.c[0] / s, a.c[1] / s, a.c[2] / s); }
static double vdot(V3 a, V3 b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; } /* vec3d.cc:35 */
static double vmag(V3 a) { return sqrt(a.c[0] * a.c[0] + a.c[1] * a.c[1] + a.c[2] * a.c[2]); } /* vec3d.cc:39 */
static V3 vnorm(V3 a) { return vdiv(a, vmag(a)); }                                   /* vec3d.cc:51 */
static V3 vcross(V3 a, V3 b) {                                                        /* vec3d.cc:43 */
    return v3(a.c[1] * b.c[2] - a.c[2] * b.c[1], -a.c[0] * b.c[2] + a.c[2] * b.c[0], a.c[0] * b.c[1] - a.c[1] * b.c[0]);
}
static V3 vzero(void) { return v3(0., 0., 0.); }

static double dmin(double a, double b) { return (b < a) ? b : a; }   /* std::min */
static double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

/* growable int list */
typedef struct { int n, cap; int* a; } ivec;
static void iv_push(ivec* v, int x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = (int*)realloc(v->a, sizeof(int) * v->cap); }
    v->a[v->n++] = x;
}
static int iv_has(const ivec* v, int x) { for (int i = 0; i < v->n; ++i) if (v->a[i] == x) return 1; return 0; }
static void iv_free(ivec* v) { free(v->a); v->a = NULL; v->n = v->cap = 0; }

/* ordered map int -> V3 (std::map<int,Vec3D> semantics) */
#define MAPMAX 512
typedef struct { int n; int key[MAPMAX]; V3 val[MAPMAX]; } vmap;
static int vm_find(const vmap* m, int k) { for (int i = 0; i < m->n; ++i) if (m->key[i] == k) return i; return -1; }
static int vm_insert(vmap* m, int k, V3 v) { /* std::map::insert: no-op if present */
    if (vm_find(m, k) >= 0) return 0;
    if (m->n >= MAPMAX) return -1;
    int i = m->n;
    while (i > 0 && m->key[i - 1] > k) { m->key[i] = m->key[i - 1]; m->val[i] = m->val[i - 1]; --i; }
    m->key[i] = k; m->val[i] = v; m->n++;
    return 0;
}

/* ------------------------------------------------------------------ */
/* context                                                            */
/* ------------------------------------------------------------------ */
#define NVM 6   /* internal stride of per-cell variable arrays: p,u,v,w,T (+ nu~ with SA) */
typedef struct {   /* NS_Cell_State (ns_state_cache.h:26-48) + SA nu~ */
    double p, p_center, T, T_center, rho, a, H, volume;
    V3 V, V_center, Vn;
    double update[NVM];
    double nut, nut_center;
} cstate;
typedef struct {   /* NS_Face_State (ns_state_cache.h:50-59) + SA face quantities */
    int index, parent, neighbor;
    double T, mu, lambda;
    V3 V, gradu, gradv, gradw, gradT, normal, tangent1, tangent2, left2right;
    double area;
    int bc;
    V3 gradn;
    double nut, rho, mu_t;
} fstate;

struct orc_ctx {
    fcfd_options opt;
    int dim;
    double R;               /* gas constant UNIV_GAS_CONST/Mw (material.h:26) */
    double Cp;              /* ideal gas gamma*R/(gamma-1) (material.cc:142-145) */
    /* nodes */
    int nnode; V3* node; ivec* node_cells;
    /* cells: [0,ncell) interior, [ncell,ntot) boundary ghosts */
    int ncell, ntot, ncap;
    ivec* cnodes; ivec* cfaces; ivec* cnbr;
    V3* cc; double* vol; double* lscale; double* wdist; int* cbc;
    /* faces */
    int nface, fcap;
    int *fparent, *fneigh, *fbc, *fsym, *face2bc;
    ivec* fnodes;
    V3 *fc, *fn;
    double *farea, *fwd, *fdiss;
    double glscale, total_volume;
    /* BC node flags and areas */
    int nbc; unsigned char** bcflag; double* bc_area; double* bc_total_area;
    /* interpolation + gradient maps (CSR, ascending keys) */
    int *avg_ptr, *avg_idx; double* avg_w;
    int *navg_ptr, *navg_idx; double* navg_w;   /* node averages (node_interpolation_weights) */
    int *g_ptr, *g_idx; V3* g_w;
    /* state */
    int nv;                         /* 5, or 6 with the SA model */
    double *p, *u, *v, *w, *T;      /* ntot */
    double* nut;                    /* ntot (SA) */
    V3 *gp, *gu, *gv, *gw, *gT;     /* ntot */
    V3* gn;                         /* ntot (SA) */
    double* phi;                    /* ntot*NVM */
    double* upd;                    /* ncell*NVM */
    double *Rs, *Rabs;              /* ncell*NVM */
    double *fC, *fD;                /* F*NVM: convective and diffusive flux of each face (last assemble) */
    double *JL, *JR, *Dg;           /* F*NVM^2 [var][eq], F*NVM^2, N*NVM^2 [row][col] */
    double* loads;                  /* nbc*6 */
    double* wL;                     /* F: weightL of each face's convective flux (last orc_assemble) */
    /* segregated two-equation RANS (src/rans/): k, omega, mu_t and the
     * strain rate of every cell (ghosts included), the gradients of k and
     * omega (interior cells; the ghosts' stay 0, rans.cc:172-178 never
     * writes them), the inlet values of rans_set_bcs.cc and which BCs fix
     * k and omega (inlets) and mu_t (inlets, no-slip walls) */
    int kw;                         /* opt.turbulence == FCFD_TURB_KW */
    double *kk, *om, *mut, *strain; /* ntot */
    V3 *gk, *gom;                   /* ntot */
    double *kw_kbc, *kw_obc, *kw_mbc;   /* nbc */
    int *kw_kfix, *kw_mfix;             /* nbc */
    double kw_lam_last;             /* laminar viscosity of the face loop's last face (terms() reads it in its cell loop) */
    int err; char msg[256];
};

static char g_err[256];
static void seterr(orc_ctx* o, int code, const char* m) {
    if (o) { o->err = code; snprintf(o->msg, sizeof o->msg, "%s", m); }
    snprintf(g_err, sizeof g_err, "%s", m);
}
const char* orc_last_error(const orc_ctx* o) { return o ? o->msg : g_err; }

/* ------------------------------------------------------------------ */
/* material (src/material/material.cc:108-153)                        */
/* ------------------------------------------------------------------ */
static double m_rho(const orc_ctx* o, double p, double T) { return (p + o->opt.Pref) / (o->R * (T + o->opt.Tref)); }
static double m_p(const orc_ctx* o, double rho, double T) { return rho * o->R * (T + o->opt.Tref) - o->opt.Pref; }
static double m_T(const orc_ctx* o, double p, double rho) { return (p + o->opt.Pref) / (o->R * rho) - o->opt.Tref; }
static double m_a(const orc_ctx* o, double p, double T) { return sqrt(o->opt.gamma * (p + o->opt.Pref) / m_rho(o, p, T)); }
static double m_visc(const orc_ctx* o, double T) {
    if (o->opt.visc_model == FCFD_VISC_SUTHERLAND)
        return o->opt.sut_mu_ref * pow((T + o->opt.Tref) / o->opt.sut_T_ref, 1.5) * (o->opt.sut_T_ref + o->opt.sut_S) / (T + o->opt.Tref + o->opt.sut_S);
    return o->opt.mu;
}
static double m_cond(const orc_ctx* o, double T) {
    if (o->opt.cond_model == FCFD_COND_PRANDTL) return o->Cp * m_visc(o, T) / o->opt.Pr;
    return o->opt.lambda;
}

/* ------------------------------------------------------------------ */
/* Spalart-Allmaras model (build-only; the reference has none).        */
/* SA without ft2, with the S~ modification and the negative-nu~ branch */
/* of Allmaras, Johnson & Spalart (ICCFD7-1902, 2012), compressible    */
/* form without the grad(rho).grad(nu~) term:                           */
/*   d(rho nu~)/dt + div(rho V nu~) = rho (P - D)                       */
/*       + (1/sigma) [div((mu + rho nu~ fn) grad nu~) + cb2 rho |grad nu~|^2] */
/* ------------------------------------------------------------------ */
static const double sa_cb1 = 0.1355, sa_sigma = 2. / 3., sa_cb2 = 0.622, sa_kappa = 0.41, sa_cw2 = 0.3, sa_cw3 = 2.,
                    sa_cv1 = 7.1, sa_ct3 = 1.2, sa_cv2 = 0.7, sa_cv3 = 0.9, sa_cn1 = 16.;
/* scale of nu~ for the FD step: the largest |nu~| prescribed at a BC, 1 if none */
static double sa_nut_scale(const orc_ctx* o) {
    double s = 0.;
    for (int b = 0; b < o->nbc; ++b) s = dmax(s, fabs(o->opt.bc[b].nut));
    return s > 0. ? s : 1.;
}
static double sa_cw1(void) { return sa_cb1 / (sa_kappa * sa_kappa) + (1. + sa_cb2) / sa_sigma; }
/* fv1 = chi^3 / (chi^3 + cv1^3) */
static double sa_fv1(double chi) { double c3 = chi * chi * chi; return c3 / (c3 + sa_cv1 * sa_cv1 * sa_cv1); }
/* eddy viscosity mu_t = rho nu~ fv1 (nu~ >= 0), 0 otherwise */
static double sa_mut(double rho, double nut, double mu) {
    if (!(nut > 0.)) return 0.;
    double chi = rho * nut / mu;
    return rho * nut * sa_fv1(chi);
}
/* diffusion coefficient factor fn (1 for nu~ >= 0) */
static double sa_fn(double rho, double nut, double mu) {
    if (nut >= 0.) return 1.;
    double chi = rho * nut / mu;
    double c3 = chi * chi * chi;
    return (sa_cn1 + c3) / (sa_cn1 - c3);
}
/* source per unit volume: rho (P - D) + (cb2/sigma) rho |grad nu~|^2, from
 * the cell state, the vorticity magnitude of the cell velocity gradient and
 * the wall distance d */
static double sa_source(double rho, double mu, double nut, V3 gu, V3 gv, V3 gw, V3 gn, double d, double* prod, double* dest) {
    double wx = gw.c[1] - gv.c[2], wy = gu.c[2] - gw.c[0], wz = gv.c[0] - gu.c[1];
    double Om = sqrt(wx * wx + wy * wy + wz * wz);
    double P, D;
    if (nut >= 0.) {
        double chi = rho * nut / mu;
        double fv1 = sa_fv1(chi);
        double fv2 = 1. - chi / (1. + chi * fv1);
        double k2d2 = sa_kappa * sa_kappa * d * d;
        double Sbar = nut * fv2 / k2d2;
        double St;
        if (Sbar >= -sa_cv2 * Om) St = Om + Sbar;
        else St = Om + Om * (sa_cv2 * sa_cv2 * Om + sa_cv3 * Sbar) / ((sa_cv3 - 2. * sa_cv2) * Om - Sbar);
        double r = (St > 0.) ? dmin(nut / (St * k2d2), 10.) : 10.;
        double r2 = r * r;
        double g = r + sa_cw2 * (r2 * r2 * r2 - r);
        double g2 = g * g, c2 = sa_cw3 * sa_cw3;
        double c6 = c2 * c2 * c2;
        double fw = g * pow((1. + c6) / (g2 * g2 * g2 + c6), 1. / 6.);
        P = sa_cb1 * St * nut;
        D = sa_cw1() * fw * (nut / d) * (nut / d);
    } else {
        P = sa_cb1 * (1. - sa_ct3) * Om * nut;
        D = -1. * sa_cw1() * (nut / d) * (nut / d);
    }
    if (prod) *prod = P;
    if (dest) *dest = D;
    return rho * (P - D) + sa_cb2 / sa_sigma * rho * vdot(gn, gn);
}

/* ------------------------------------------------------------------ */
/* Grid::create_nodes_cells (src/grid/grid_create_elements.cc:27-182) */
/* ------------------------------------------------------------------ */
static int create_nodes_cells(orc_ctx* o, const fcfd_mesh* m) {
    int* g2l = (int*)malloc(sizeof(int) * (size_t)m->n_nodes);
    for (int i = 0; i < m->n_nodes; ++i) g2l[i] = -1;
    o->ncell = m->n_cells;
    o->ncap = m->n_cells + 1;
    o->cnodes = (ivec*)calloc(o->ncap, sizeof(ivec));
    o->cfaces = (ivec*)calloc(o->ncap, sizeof(ivec));
    o->node = (V3*)malloc(sizeof(V3) * (size_t)m->n_nodes);
    o->nnode = 0;
    for (int c = 0; c < m->n_cells; ++c) {                       /* :35-103 */
        int nn = m->cell_ptr[c + 1] - m->cell_ptr[c];
        if (nn != 4 && nn != 5 && nn != 6 && nn != 8) { free(g2l); seterr(o, FCFD_EMESH, "unsupported cell node count"); return FCFD_EMESH; }
        for (int n = 0; n < nn; ++n) {
            int ng = m->cell_nodes[m->cell_ptr[c] + n];
            if (ng < 0 || ng >= m->n_nodes) { free(g2l); seterr(o, FCFD_EINVAL, "node id out of range"); return FCFD_EINVAL; }
            if (g2l[ng] < 0) {
                g2l[ng] = o->nnode;
                o->node[o->nnode] = v3(m->xyz[3 * ng], m->xyz[3 * ng + 1], m->xyz[3 * ng + 2]);
                o->nnode++;
            }
            iv_push(&o->cnodes[c], g2l[ng]);
        }
        int nf = (nn == 4) ? 4 : (nn == 5) ? 5 : (nn == 6) ? 5 : 6; /* :68-81 */
        for (int i = 0; i < nf; ++i) iv_push(&o->cfaces[c], -1);
    }
    /* node -> cells (:126-142) */
    o->node_cells = (ivec*)calloc(o->nnode, sizeof(ivec));
    for (int c = 0; c < o->ncell; ++c)
        for (int cn = 0; cn < o->cnodes[c].n; ++cn) {
            int n = o->cnodes[c].a[cn];
            if (!iv_has(&o->node_cells[n], c)) iv_push(&o->node_cells[n], c);
        }
    /* cell -> node-sharing neighbours incl. self (:146-164) */
    o->cnbr = (ivec*)calloc(o->ncap, sizeof(ivec));
    for (int c = 0; c < o->ncell; ++c)
        for (int cn = 0; cn < o->cnodes[c].n; ++cn) {
            int n = o->cnodes[c].a[cn];
            for (int nc = 0; nc < o->node_cells[n].n; ++nc) {
                int c2 = o->node_cells[n].a[nc];
                if (!iv_has(&o->cnbr[c], c2)) iv_push(&o->cnbr[c], c2);
            }
        }
    /* BC node sets in local numbering (:168-178) */
    o->nbc = m->n_bc;
    o->bcflag = (unsigned char**)calloc(o->nbc > 0 ? o->nbc : 1, sizeof(unsigned char*));
    for (int b = 0; b < o->nbc; ++b) {
        o->bcflag[b] = (unsigned char*)calloc(o->nnode > 0 ? o->nnode : 1, 1);
        for (int k = m->bc_ptr[b]; k < m->bc_ptr[b + 1]; ++k) {
            int ng = m->bc_nodes[k];
            if (ng >= 0 && ng < m->n_nodes && g2l[ng] >= 0) o->bcflag[b][g2l[ng]] = 1;
        }
    }
    free(g2l);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Grid::create_faces (grid_create_elements.cc:184-442)               */
/* ------------------------------------------------------------------ */
static const int hexaFaces[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {1, 2, 6, 5}, {0, 4, 7, 3}, {1, 5, 4, 0}, {2, 3, 7, 6}};
static const int prismFaces[5][4] = {{0, 2, 1, 0}, {3, 4, 5, 0}, {0, 3, 5, 2}, {1, 2, 5, 4}, {0, 1, 4, 3}};
static const int pyraFaces[5][4] = {{0, 3, 2, 1}, {0, 1, 4, 0}, {1, 2, 4, 0}, {3, 4, 2, 0}, {0, 4, 3, 0}};
static const int tetraFaces[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};

static int have_nodes(const orc_ctx* o, int c, int nl, const int* list) {   /* :978-986 */
    for (int i = 0; i < nl; ++i) {
        int match = 0;
        for (int j = 0; j < o->cnodes[c].n; ++j) if (list[i] == o->cnodes[c].a[j]) { match = 1; break; }
        if (!match) return 0;
    }
    return 1;
}

static void face_grow(orc_ctx* o) {
    if (o->nface < o->fcap) return;
    int nc = o->fcap ? 2 * o->fcap : 1024;
    o->fparent = (int*)realloc(o->fparent, sizeof(int) * nc);
    o->fneigh = (int*)realloc(o->fneigh, sizeof(int) * nc);
    o->fbc = (int*)realloc(o->fbc, sizeof(int) * nc);
    o->fnodes = (ivec*)realloc(o->fnodes, sizeof(ivec) * nc);
    memset(o->fnodes + o->fcap, 0, sizeof(ivec) * (nc - o->fcap));
    o->fcap = nc;
}

static int create_faces(orc_ctx* o) {
    o->nface = 0;
    unsigned char* repeated_cell = (unsigned char*)calloc(o->ncell > 0 ? o->ncell : 1, 1);
    for (int c = 0; c < o->ncell; ++c) {
        int nn = o->cnodes[c].n;
        int degenerate_face_count = 0;
        for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
            int fnn;
            switch (nn) {
            case 4: fnn = 3; break;
            case 5: fnn = (cf < 1) ? 4 : 3; break;
            case 6: fnn = (cf < 2) ? 3 : 4; break;
            default: fnn = 4; break;
            }
            int tn[4];
            for (int fn = 0; fn < fnn; ++fn) {
                switch (nn) {
                case 4: tn[fn] = o->cnodes[c].a[tetraFaces[cf][fn]]; break;
                case 5: tn[fn] = o->cnodes[c].a[pyraFaces[cf][fn]]; break;
                case 6: tn[fn] = o->cnodes[c].a[prismFaces[cf][fn]]; break;
                default: tn[fn] = o->cnodes[c].a[hexaFaces[cf][fn]]; break;
                }
            }
            /* repeated nodes (:274-294): keep the first appearance of each; a
             * face left with two nodes is degenerate and not created */
            int un[4], nu = 0, degenerate = 0;
            for (int fn = 0; fn < fnn; ++fn) {
                int skip = 0;
                for (int i = 0; i < fn; ++i) if (tn[fn] == tn[i]) { skip = 1; break; }
                if (!skip) un[nu++] = tn[fn];
            }
            if (nu != fnn) {
                repeated_cell[c] = 1;
                if (nu == 2) { degenerate = 1; degenerate_face_count++; }
                fnn = nu;
                for (int fn = 0; fn < fnn; ++fn) tn[fn] = un[fn];
            }
            int internal = 0, unique = 1, neighbor = -1;              /* :296-314 */
            int n0 = tn[0];
            for (int nc = 0; nc < o->node_cells[n0].n; ++nc) {
                int i = o->node_cells[n0].a[nc];
                if (i != c && i < o->ncell && have_nodes(o, i, fnn, tn)) {
                    if (i > c) { neighbor = i; internal = 1; }
                    else unique = 0;
                }
            }
            if (!unique || degenerate) continue;
            int bcv = -1;  /* INTERNAL_FACE */
            if (!internal) {                                          /* :318-363 */
                bcv = -2;  /* UNASSIGNED_FACE */
                int matched[FCFD_MAX_BC], nmatched = 0, cell_matched = -1;
                for (int nbc = 0; nbc < o->nbc; ++nbc) {
                    int match = 1;
                    for (int i = 0; i < fnn; ++i) if (!o->bcflag[nbc][tn[i]]) { match = 0; break; }
                    if (match && nmatched < FCFD_MAX_BC) matched[nmatched++] = nbc;
                    if (cell_matched == -1) {
                        match = 1;
                        for (int i = 0; i < nn; ++i) if (!o->bcflag[nbc][o->cnodes[c].a[i]]) { match = 0; break; }
                        if (match) cell_matched = nbc;
                    }
                }
                if (nmatched > 1) {
                    for (int k = 0; k < nmatched; ++k) if (matched[k] != cell_matched) { bcv = matched[k]; break; }
                } else if (nmatched == 1) bcv = matched[0];
            }
            face_grow(o);
            int f = o->nface;
            for (int i = 0; i < o->cfaces[c].n; ++i) if (o->cfaces[c].a[i] < 0) { o->cfaces[c].a[i] = f; break; }  /* :364-369 */
            if (internal)
                for (int i = 0; i < o->cfaces[neighbor].n; ++i) if (o->cfaces[neighbor].a[i] < 0) { o->cfaces[neighbor].a[i] = f; break; }
            o->fparent[f] = c; o->fneigh[f] = neighbor; o->fbc[f] = bcv;
            o->fnodes[f].n = 0;
            for (int i = 0; i < fnn; ++i) iv_push(&o->fnodes[f], tn[i]);
            o->nface++;
        }
        o->cfaces[c].n -= degenerate_face_count;                      /* :424 */
    }
    /* cells with repeated nodes (:426-467): a hexahedron with two repeated
     * nodes becomes a prism whose nodes are read off its triangular faces,
     * each rotated to start at a repeated node, in the cell's face order */
    for (int c = 0; c < o->ncell; ++c) {
        if (!repeated_cell[c]) continue;
        int rep[64], nrep = 0;
        for (int cn = 0; cn < o->cnodes[c].n; ++cn)
            for (int cn2 = 0; cn2 < cn; ++cn2)
                if (o->cnodes[c].a[cn] == o->cnodes[c].a[cn2] && nrep < 64) rep[nrep++] = o->cnodes[c].a[cn];
        if (o->cnodes[c].n != 8 || nrep != 2) continue;
        o->cnodes[c].n = 0;
        for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
            const ivec* fnod = &o->fnodes[o->cfaces[c].a[cf]];
            if (fnod->n != 3) continue;
            int fn;
            for (fn = 0; fn < 3; ++fn) {
                if (fnod->a[fn] == rep[0]) break;
                if (fnod->a[fn] == rep[1]) break;
            }
            if (fn == 3) continue;
            for (int k = 0; k < 3; ++k) iv_push(&o->cnodes[c], fnod->a[(fn + k) % 3]);
        }
    }
    free(repeated_cell);
    for (int f = 0; f < o->nface; ++f) if (o->fbc[f] == -2) { seterr(o, FCFD_EBC, "boundary condition could not be found for a face"); return FCFD_EBC; } /* set_bcs.cc:147-150 */
    return 0;
}

/* ------------------------------------------------------------------ */
/* Grid::areas_volumes (src/grid/grid.cc:123-178)                      */
/* ------------------------------------------------------------------ */
static void areas_volumes(orc_ctx* o) {
    o->fc = (V3*)malloc(sizeof(V3) * o->nface);
    o->fn = (V3*)malloc(sizeof(V3) * o->nface);
    o->farea = (double*)malloc(sizeof(double) * o->nface);
    for (int f = 0; f < o->nface; ++f) {
        ivec* fnod = &o->fnodes[f];
        if (fnod->n == 4) {
            V3 d1 = vsub(o->node[fnod->a[2]], o->node[fnod->a[0]]);
            V3 d2 = vsub(o->node[fnod->a[3]], o->node[fnod->a[1]]);
            V3 nrm = vcross(d1, d2);
            o->farea[f] = 0.5 * vmag(nrm);
            o->fn[f] = vnorm(nrm);
        } else {
            V3 areaVec = vzero();
            V3 cen = o->node[fnod->a[0]];
            for (int n = 1; n < fnod->n; ++n) {
                int next = n + 1;
                if (next == fnod->n) next = 0;
                V3 patch = vsc(0.5, vcross(vsub(o->node[fnod->a[n]], cen), vsub(o->node[fnod->a[next]], cen)));
                areaVec = vadd(areaVec, patch);
            }
            o->farea[f] = vmag(areaVec);
            o->fn[f] = vnorm(areaVec);
        }
        V3 s = vzero();
        for (int n = 0; n < fnod->n; ++n) s = vadd(s, o->node[fnod->a[n]]);
        o->fc[f] = vdiv(s, (double)fnod->n);
    }
    o->cc = (V3*)malloc(sizeof(V3) * o->ncap);
    o->vol = (double*)malloc(sizeof(double) * o->ncap);
    double total = 0.;
    for (int c = 0; c < o->ncell; ++c) {
        V3 s = vzero();
        for (int cn = 0; cn < o->cnodes[c].n; ++cn) s = vadd(s, o->node[o->cnodes[c].a[cn]]);
        o->cc[c] = vdiv(s, (double)o->cnodes[c].n);
        o->vol[c] = 0.;
        for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
            int f = o->cfaces[c].a[cf];
            V3 base = vsc(o->farea[f], o->fn[f]);
            if (o->fparent[f] == c) base = vmul(base, -1.);
            V3 bn = vnorm(base);
            V3 height = vsc(vdot(vsub(o->cc[c], o->fc[f]), bn), bn);
            o->vol[c] += vdot(base, height) / 3.;
        }
        total += o->vol[c];
    }
    o->total_volume = total;
}

/* ------------------------------------------------------------------ */
/* Grid::create_boundary_ghosts (grid_create_elements.cc:702-772)      */
/* ------------------------------------------------------------------ */
static void create_boundary_ghosts(orc_ctx* o) {
    int nb = 0;
    for (int f = 0; f < o->nface; ++f) if (o->fbc[f] >= 0) nb++;
    int ntot = o->ncell + nb;
    o->ncap = ntot + 1;
    o->cnodes = (ivec*)realloc(o->cnodes, sizeof(ivec) * o->ncap);
    o->cfaces = (ivec*)realloc(o->cfaces, sizeof(ivec) * o->ncap);
    o->cnbr = (ivec*)realloc(o->cnbr, sizeof(ivec) * o->ncap);
    memset(o->cnodes + o->ncell, 0, sizeof(ivec) * (o->ncap - o->ncell));
    memset(o->cfaces + o->ncell, 0, sizeof(ivec) * (o->ncap - o->ncell));
    memset(o->cnbr + o->ncell, 0, sizeof(ivec) * (o->ncap - o->ncell));
    o->cc = (V3*)realloc(o->cc, sizeof(V3) * o->ncap);
    o->vol = (double*)realloc(o->vol, sizeof(double) * o->ncap);
    o->cbc = (int*)malloc(sizeof(int) * o->ncap);
    for (int c = 0; c < o->ncell; ++c) o->cbc[c] = -1;
    int g = o->ncell;
    for (int b = 0; b < o->nbc; ++b)
        for (int f = 0; f < o->nface; ++f)
            if (o->fbc[f] == b) {
                int par = o->fparent[f];
                o->cbc[g] = b;
                o->vol[g] = o->vol[par];
                o->cc[g] = vadd(o->cc[par], vsc(2. * vdot(vsub(o->fc[f], o->cc[par]), o->fn[f]), o->fn[f])); /* :725-726 */
                o->fneigh[f] = g;
                for (int fn = 0; fn < o->fnodes[f].n; ++fn) iv_push(&o->node_cells[o->fnodes[f].a[fn]], g);
                g++;
            }
    o->ntot = g;
    for (int c = 0; c < o->ncell; ++c)                                /* :746-769 */
        for (int cn = 0; cn < o->cnodes[c].n; ++cn) {
            int n = o->cnodes[c].a[cn];
            for (int ng = 0; ng < o->node_cells[n].n; ++ng) {
                int gg = o->node_cells[n].a[ng];
                if (gg >= o->ncell && !iv_has(&o->cnbr[c], gg)) {
                    iv_push(&o->cnbr[c], gg);
                    iv_push(&o->cnbr[gg], c);
                }
            }
        }
}

/* ------------------------------------------------------------------ */
/* set_bcs (src/set_bcs.cc:35-260), BC regions = mesh node sets       */
/* ------------------------------------------------------------------ */
/* closest wall face: the brute-force scan (set_bcs.cc:178-249): the
 * smallest |x - x_w| over the wall face centroids, the first in face order
 * on ties (*who; -1 without walls) */
static int g_wall_brute = 0;
static double wall_brute(const V3* wc, int n, V3 x, int* who) {
    double best = 1.e20;
    *who = -1;
    for (int s = 0; s < n; ++s) {
        double d = vmag(vsub(x, wc[s]));
        if (best > d) { best = d; *who = s; }
    }
    return best;
}

/* uniform buckets over the wall centroids' bounding box, about two
 * centroids per bucket, each bucket's centroids in face order */
typedef struct { int n[3]; double lo[3], hi[3], h[3]; int *ptr, *idx; } wallgrid_t;

static void wallgrid_build(wallgrid_t* g, const V3* wc, int nw) {
    for (int a = 0; a < 3; ++a) { g->lo[a] = 1.e300; g->hi[a] = -1.e300; }
    for (int s = 0; s < nw; ++s)
        for (int a = 0; a < 3; ++a) { g->lo[a] = dmin(g->lo[a], wc[s].c[a]); g->hi[a] = dmax(g->hi[a], wc[s].c[a]); }
    double ext[3], vol = 1.;
    int nd = 0;
    for (int a = 0; a < 3; ++a) { ext[a] = nw ? g->hi[a] - g->lo[a] : 0.; if (ext[a] > 0.) { vol *= ext[a]; nd++; } }
    double target = nw > 2 ? nw / 2.0 : 1.0, side = nd ? pow(vol / target, 1.0 / nd) : 1.0;
    int total = 1;
    for (int a = 0; a < 3; ++a) {
        g->n[a] = (ext[a] > 0. && side > 0.) ? (int)ceil(ext[a] / side) : 1;
        if (g->n[a] < 1) g->n[a] = 1;
        if (g->n[a] > 1024) g->n[a] = 1024;
        g->h[a] = ext[a] > 0. ? ext[a] / g->n[a] : 1.;
        total *= g->n[a];
    }
    g->ptr = (int*)calloc(total + 1, sizeof(int));
    g->idx = (int*)malloc(sizeof(int) * (nw + 1));
    int* cell = (int*)malloc(sizeof(int) * (nw + 1));
    for (int s = 0; s < nw; ++s) {
        int b[3];
        for (int a = 0; a < 3; ++a) {
            b[a] = (int)((wc[s].c[a] - g->lo[a]) / g->h[a]);
            if (b[a] < 0) b[a] = 0;
            if (b[a] >= g->n[a]) b[a] = g->n[a] - 1;
        }
        cell[s] = (b[2] * g->n[1] + b[1]) * g->n[0] + b[0];
        g->ptr[cell[s] + 1]++;
    }
    for (int k = 0; k < total; ++k) g->ptr[k + 1] += g->ptr[k];
    int* fill = (int*)malloc(sizeof(int) * (total + 1));
    memcpy(fill, g->ptr, sizeof(int) * (total + 1));
    for (int s = 0; s < nw; ++s) g->idx[fill[cell[s]]++] = s;   /* ascending face order per bucket */
    free(fill); free(cell);
}

static void wallgrid_free(wallgrid_t* g) { free(g->ptr); free(g->idx); }

/* the brute force's result by shells of buckets around the query's bucket:
 * a bucket outside the searched box along axis a is at least
 * sqrt(slab_a^2 + sum_{b != a} gap_b^2) away (slab: the distance to the box's
 * side; gap: to the grid's extent), so the search stops once the best
 * distance is below every such bound */
static double wallgrid_nearest(const wallgrid_t* g, const V3* wc, V3 x, int* who) {
    double best = 1.e20;
    int bi = -1;
    *who = -1;
    if (g->ptr[g->n[0] * g->n[1] * g->n[2]] == 0) return best;
    int q[3];
    double gap[3];
    for (int a = 0; a < 3; ++a) {
        q[a] = (int)floor((x.c[a] - g->lo[a]) / g->h[a]);
        if (q[a] < 0) q[a] = 0;
        if (q[a] >= g->n[a]) q[a] = g->n[a] - 1;
        gap[a] = dmax(0., dmax(g->lo[a] - x.c[a], x.c[a] - g->hi[a]));
    }
    for (int r = 0;; ++r) {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = q[a] - r; hi[a] = q[a] + r; if (lo[a] < 0) lo[a] = 0; if (hi[a] >= g->n[a]) hi[a] = g->n[a] - 1; }
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    /* only the new shell's buckets */
                    if (r > 0 && i > q[0] - r && i < q[0] + r && j > q[1] - r && j < q[1] + r && k > q[2] - r && k < q[2] + r) continue;
                    int b = (k * g->n[1] + j) * g->n[0] + i;
                    for (int e = g->ptr[b]; e < g->ptr[b + 1]; ++e) {
                        int s = g->idx[e];
                        double d = vmag(vsub(x, wc[s]));
                        if (d < best || (d == best && s < bi)) { best = d; bi = s; }
                    }
                }
        /* lower bound on every bucket beyond the box */
        double bound = 1.e300;
        int beyond = 0;
        double g2 = gap[0] * gap[0] + gap[1] * gap[1] + gap[2] * gap[2];
        for (int a = 0; a < 3; ++a) {
            double others = g2 - gap[a] * gap[a];
            if (q[a] - r > 0) {
                double slab = dmax(0., x.c[a] - (g->lo[a] + (q[a] - r) * g->h[a]));
                bound = dmin(bound, sqrt(slab * slab + others));
                beyond = 1;
            }
            if (q[a] + r < g->n[a] - 1) {
                double slab = dmax(0., (g->lo[a] + (q[a] + r + 1) * g->h[a]) - x.c[a]);
                bound = dmin(bound, sqrt(slab * slab + others));
                beyond = 1;
            }
        }
        if (!beyond || best < bound) break;
    }
    *who = bi;
    return best;
}

/* the tests switch the exact bucketed search to the reference's scan */
int orc_set_wall_brute_force(int on) { g_wall_brute = on; return 0; }

static void set_bcs_grid(orc_ctx* o) {
    o->fsym = (int*)calloc(o->nface, sizeof(int));
    o->face2bc = (int*)calloc(o->nface, sizeof(int));
    o->bc_area = (double*)calloc(o->nbc + 1, sizeof(double));
    o->bc_total_area = (double*)calloc(o->nbc + 1, sizeof(double));
    for (int b = 0; b < o->nbc; ++b) {                                /* :86-98 */
        int counter = 0;
        double area = 0.;
        for (int f = 0; f < o->nface; ++f)
            if (o->fbc[f] == b) {
                o->fsym[f] = 0;
                o->face2bc[f] = counter++;
                area += o->farea[f];
                if (o->opt.bc[b].type == FCFD_BC_SYMMETRY) o->fsym[f] = 1;
            }
        o->bc_area[b] = area;
        o->bc_total_area[b] = area;
    }
    for (int f = 0; f < o->nface; ++f) {                              /* :146-161 */
        int b = o->fbc[f];
        if (b >= 0 && (o->opt.bc[b].type == FCFD_BC_INLET || o->opt.bc[b].type == FCFD_BC_OUTLET)) {
            int nb = o->fneigh[f], par = o->fparent[f];
            o->cc[nb] = vadd(o->cc[par], vsc(2., vsub(o->fc[f], o->cc[par])));
        }
    }
    /* closest wall distance, brute force (:178-249) */
    int nsf = 0;
    for (int f = 0; f < o->nface; ++f) {
        int b = o->fbc[f];
        if (b >= 0 && o->opt.bc[b].type == FCFD_BC_WALL && o->opt.bc[b].kind != FCFD_KIND_SLIP) nsf++;
    }
    V3* wc = (V3*)malloc(sizeof(V3) * (nsf + 1));
    V3* wn = (V3*)malloc(sizeof(V3) * (nsf + 1));
    int k = 0;
    for (int f = 0; f < o->nface; ++f) {
        int b = o->fbc[f];
        if (b >= 0 && o->opt.bc[b].type == FCFD_BC_WALL && o->opt.bc[b].kind != FCFD_KIND_SLIP) { wc[k] = o->fc[f]; wn[k] = o->fn[f]; k++; }
    }
    /* The reference scans every wall face for every cell (and face); the
     * minimum of the same distances, with the first closest face in face
     * order on exact ties, is found here by a bucketed search, so the
     * results are the brute force's bit for bit (brute_force = 1 runs the
     * scan itself, for the tests) */
    wallgrid_t wg;
    wallgrid_build(&wg, wc, nsf);
    o->wdist = (double*)malloc(sizeof(double) * o->ncap);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int c = 0; c < o->ntot; ++c) {
        int who;
        o->wdist[c] = g_wall_brute ? wall_brute(wc, nsf, o->cc[c], &who) : wallgrid_nearest(&wg, wc, o->cc[c], &who);
    }
    o->fwd = (double*)malloc(sizeof(double) * o->nface);
    o->fdiss = (double*)malloc(sizeof(double) * o->nface);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int f = 0; f < o->nface; ++f) {
        int b = o->fbc[f];
        o->fwd[f] = 1.e20;
        if (b < 0 || o->opt.bc[b].type != FCFD_BC_WALL || o->opt.bc[b].kind == FCFD_KIND_SLIP) {
            int who = -1;
            o->fwd[f] = g_wall_brute ? wall_brute(wc, nsf, o->fc[f], &who) : wallgrid_nearest(&wg, wc, o->fc[f], &who);
            V3 fN = who >= 0 ? wn[who] : vzero();
            o->fdiss[f] = 1. - fabs(fN.c[0] * o->fn[f].c[0] + fN.c[1] * o->fn[f].c[1] + fN.c[2] * o->fn[f].c[2]);
        } else { o->fwd[f] = 0.; o->fdiss[f] = 0.; }
    }
    wallgrid_free(&wg);
    free(wc); free(wn);
}

/* set_lengthScales (src/main.cc:456-491) */
static void set_length_scales(orc_ctx* o) {
    o->lscale = (double*)malloc(sizeof(double) * o->ncap);
    for (int c = 0; c < o->ncell; ++c) {
        o->lscale[c] = 1.e20;
        for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
            int f = o->cfaces[c].a[cf];
            double h = fabs(vdot(o->fn[f], vsub(o->fc[f], o->cc[c])));
            int skip = (o->fbc[f] >= 0 && o->opt.bc[o->fbc[f]].type == FCFD_BC_SYMMETRY);
            if (!skip) o->lscale[c] = dmin(o->lscale[c], h);
        }
    }
    for (int c = o->ncell; c < o->ntot; ++c) o->lscale[c] = 0.;
    if (o->dim == 3) o->glscale = pow(o->total_volume, 1. / 3.);
    else {
        o->glscale = 0.;
        for (int b = 0; b < o->nbc; ++b)
            if (o->opt.bc[b].type == FCFD_BC_SYMMETRY) o->glscale = dmax(o->glscale, sqrt(0.5 * o->bc_total_area[b]));
    }
}

/* ------------------------------------------------------------------ */
/* face_interpolation_weights (src/face_interpolation_weights.cc:32-138)
 * + Interpolate::calculate_weights/sort_stencil/interpolate_line
 * (src/interpolate/interpolate.cc:36-121,320-343)                      */
/* ------------------------------------------------------------------ */
static int cmp_int(const void* a, const void* b) { int x = *(const int*)a, y = *(const int*)b; return (x > y) - (x < y); }

/* ------------------------------------------------------------------ */
/* Interpolate (src/interpolate/interpolate.cc:36-351) and gelimd      */
/* (src/utilities/utilities.cc:71-101)                                 */
/* ------------------------------------------------------------------ */
/* One Interpolate object: its weights3/weights4 members persist between
 * calls, and gelimd leaves its solution untouched on a singular system, so
 * a singular fit reuses the previous fit's solution (kept here). */
typedef struct { int dim, maxs; double tol; double w4[4], w3[3]; } interp_t;

#define GELIM_EPS 1e-12   /* EPS, src/utilities/utilities.h:26 */
static int gelimd(int n, double a[4][4], double* b, double* x) {   /* utilities.cc:71-101 */
    int i, j, k;
    double tmp, pvt;
    for (i = 0; i < n; i++) {
        pvt = a[i][i];
        if (fabs(pvt) < GELIM_EPS) {
            for (j = i + 1; j < n; j++) { if (fabs(pvt = a[j][i]) >= GELIM_EPS) break; }
            if (fabs(pvt) < GELIM_EPS) return 1;
            for (k = 0; k < n; ++k) { tmp = a[i][k]; a[i][k] = a[j][k]; a[j][k] = tmp; }
            tmp = b[j]; b[j] = b[i]; b[i] = tmp;
        }
        for (k = i + 1; k < n; k++) {
            tmp = a[k][i] / pvt;
            for (j = i + 1; j < n; j++) a[k][j] -= tmp * a[i][j];
            b[k] -= tmp * b[i];
        }
    }
    for (i = n - 1; i >= 0; i--) {
        x[i] = b[i];
        for (j = n - 1; j > i; j--) x[i] -= a[i][j] * x[j];
        x[i] /= a[i][i];
    }
    return 0;
}

/* sort_stencil(is_internal = true) (interpolate.cc:65-125): the first two
 * entries stay, the others follow by ascending distance to the point (an
 * equal distance goes before the earlier entry), truncated to maxs.  Both
 * callers pass is_internal = true.  Reorders ids/pts in place; returns the
 * new size.  (With a single entry the reference writes order[1] past the
 * end of its vector; that write is dropped here.) */
static int sort_stencil_keep2(int n, int maxs, V3 point, int* ids, V3* pts) {
    int m = n < maxs ? n : maxs;
    int* order = (int*)malloc(sizeof(int) * (m + 1));
    double* mind = (double*)malloc(sizeof(double) * (m + 1));
    for (int i = 0; i < m; ++i) { order[i] = 0; mind[i] = 1.e20; }
    order[0] = 0;
    if (m > 1) order[1] = 1;
    for (int d = 2; d < n; ++d) {
        double dd = vmag(vsub(pts[d], point));
        for (int q = 2; q < m; ++q) {
            if (dd <= mind[q]) {
                for (int i = m - 1; i > q; --i) order[i] = order[i - 1];
                for (int i = m - 1; i > q; --i) mind[i] = mind[i - 1];
                order[q] = d;
                mind[q] = dd;
                break;
            }
        }
    }
    int* ni = (int*)malloc(sizeof(int) * (m + 1));
    V3* np_ = (V3*)malloc(sizeof(V3) * (m + 1));
    for (int i = 0; i < m; ++i) { ni[i] = ids[order[i]]; np_[i] = pts[order[i]]; }
    for (int i = 0; i < m; ++i) { ids[i] = ni[i]; pts[i] = np_[i]; }
    free(order); free(mind); free(ni); free(np_);
    return m;
}

static int interpolate_tetra(interp_t* it, int m, const V3* P, V3 point, double* w) {   /* :127-243 */
    if (it->dim != 3 || m < 4) return 0;
    for (int i = 0; i < m; ++i) w[i] = 0.;
    int ok = 0;
    for (int s1 = 0; s1 < m - 3; ++s1)
        for (int s2 = s1 + 1; s2 < m - 2; ++s2)
            for (int s3 = s2 + 1; s3 < m - 1; ++s3)
                for (int s4 = s3 + 1; s4 < m; ++s4) {
                    V3 e1 = vsub(P[s2], P[s1]), e2 = vsub(P[s3], P[s1]), e3 = vsub(P[s3], P[s2]);
                    V3 e4 = vsub(P[s4], P[s1]), e5 = vsub(P[s4], P[s2]), e6 = vsub(P[s4], P[s3]);
                    double ave = (vmag(e1) + vmag(e2) + vmag(e3) + vmag(e4) + vmag(e5) + vmag(e6)) / 6.;
                    V3 pn = vcross(e1, e2);
                    double area = 0.5 * vmag(pn);
                    pn = vnorm(pn);
                    double volume = area * fabs(vdot(e4, pn)) / 3.;
                    double eqv = sqrt(2.) / 12. * ave * ave * ave;
                    double skew = 1. - volume / eqv;
                    if (skew > it->tol) continue;
                    V3 cen = vdiv(vadd(vadd(vadd(P[s1], P[s2]), P[s3]), P[s4]), 4.);
                    double dist = vmag(vsub(point, cen)) / ave;
                    dist = dmax(dist, 1.e-8);
                    double tw = 1. / (dist * dist);
                    double a[4][4], b[4];
                    const int s[4] = {s1, s2, s3, s4};
                    for (int q = 0; q < 4; ++q) {
                        a[0][q] = P[s[q]].c[0]; a[1][q] = P[s[q]].c[1]; a[2][q] = P[s[q]].c[2]; a[3][q] = 1.;
                    }
                    b[0] = point.c[0]; b[1] = point.c[1]; b[2] = point.c[2]; b[3] = 1.;
                    gelimd(4, a, b, it->w4);
                    double ws = 0.;
                    for (int q = 0; q < 4; ++q) ws += it->w4[q];
                    if (fabs(ws - 1.) < 1.e-8) {
                        for (int q = 0; q < 4; ++q) w[s[q]] += tw * it->w4[q];
                        ok = 1;
                    }
                }
    double ws = 0.;
    for (int i = 0; i < m; ++i) ws += w[i];
    for (int i = 0; i < m; ++i) w[i] /= ws;
    return ok;
}

static int interpolate_tri(interp_t* it, int m, const V3* P, V3 point, double* w) {   /* :245-318 */
    if (it->dim < 2 || m < 3) return 0;
    for (int i = 0; i < m; ++i) w[i] = 0.;
    int ok = 0;
    for (int s1 = 0; s1 < m - 2; ++s1)
        for (int s2 = s1 + 1; s2 < m - 1; ++s2)
            for (int s3 = s2 + 1; s3 < m; ++s3) {
                V3 e1 = vsub(P[s2], P[s1]), e2 = vsub(P[s3], P[s1]), e3 = vsub(P[s3], P[s2]);
                V3 pn = vcross(e1, e2);
                double area = 0.5 * vmag(pn);
                pn = vnorm(pn);
                double ave = (vmag(e1) + vmag(e2) + vmag(e3)) / 3.;
                double eqa = 0.433 * ave * ave;
                double skew = 1. - area / eqa;
                if (skew > it->tol) continue;
                V3 pp = vsub(point, vsc(vdot(point, pn), pn));     /* projection through the origin (:271) */
                V3 cen = vdiv(vadd(vadd(P[s1], P[s2]), P[s3]), 3.);
                double dist = vmag(vsub(pp, cen)) / ave;
                dist = dmax(dist, 1.e-2);
                double tw = 1. / (dist * dist);
                V3 b1 = vnorm(e1);
                V3 b2 = vsc(-1., vnorm(vcross(b1, pn)));
                pp = vsub(pp, P[s1]);
                double a[4][4], b[4];
                a[0][0] = 0.; a[0][1] = vdot(e1, b1); a[0][2] = vdot(e2, b1);
                a[1][0] = 0.; a[1][1] = vdot(e1, b2); a[1][2] = vdot(e2, b2);
                a[2][0] = 1.; a[2][1] = 1.; a[2][2] = 1.;
                b[0] = vdot(pp, b1); b[1] = vdot(pp, b2); b[2] = 1.;
                gelimd(3, a, b, it->w3);
                double ws = 0.;
                for (int q = 0; q < 3; ++q) ws += it->w3[q];
                if (fabs(ws - 1.) < 1.e-8) {
                    w[s1] += tw * it->w3[0]; w[s2] += tw * it->w3[1]; w[s3] += tw * it->w3[2];
                    ok = 1;
                }
            }
    double ws = 0.;
    for (int i = 0; i < m; ++i) ws += w[i];
    for (int i = 0; i < m; ++i) w[i] /= ws;
    return ok;
}

/* calculate_weights (:36-50) after the stencil is in ids/pts (n entries):
 * returns the number of weights (the stencil size for tetra/tri fits, 2 for
 * a line, 1 for a point); ids/pts are reordered as sort_stencil leaves them */
static int wtli_weights(interp_t* it, int n, V3 point, int* ids, V3* pts, double* w, int* kind) {
    int m = sort_stencil_keep2(n, it->maxs, point, ids, pts);
    for (int i = 0; i < m; ++i) w[i] = 0.;
    if (interpolate_tetra(it, m, pts, point, w)) { *kind = 4; return m; }
    if (interpolate_tri(it, m, pts, point, w)) { *kind = 3; return m; }
    if (m >= 2) {                                                 /* interpolate_line :320-343 */
        V3 e1 = vsub(pts[1], pts[0]);
        V3 b1 = vnorm(e1);
        V3 pp = vsub(point, pts[0]);
        pp = vsc(vdot(pp, b1), b1);
        w[1] = vdot(pp, b1) / vmag(e1);
        w[0] = 1. - w[1];
        *kind = 2;
        return 2;
    }
    w[0] = 1.;                                                    /* interpolate_point :345-351 */
    *kind = 1;
    return 1;
}

/* std::map<int,double>::insert of (ids[s], w[s]) for s < nw into a CSR row
 * kept in ascending keys (a repeated key keeps its first weight) */
static void map_insert_row(int** idx, double** wt, int* cap, int* nent, const int* ids, const double* w, int nw) {
    if (*nent + nw > *cap) {
        *cap = 2 * *cap + nw;
        *idx = (int*)realloc(*idx, sizeof(int) * *cap);
        *wt = (double*)realloc(*wt, sizeof(double) * *cap);
    }
    int base = *nent;
    for (int s = 0; s < nw; ++s) {
        int key = ids[s], dup = 0;
        for (int q = base; q < *nent; ++q) if ((*idx)[q] == key) { dup = 1; break; }
        if (dup) continue;
        int q = *nent;
        while (q > base && (*idx)[q - 1] > key) { (*idx)[q] = (*idx)[q - 1]; (*wt)[q] = (*wt)[q - 1]; --q; }
        (*idx)[q] = key; (*wt)[q] = w[s];
        (*nent)++;
    }
}

/* node_interpolation_weights (src/node_interpolation_weights.cc:32-118):
 * stencil = the set of cells (ghosts included) around the node; WTLI with
 * is_internal = true, IDW (distance-proportional, as for faces) or SIMPLE */
static int node_interpolation_weights(orc_ctx* o) {
    int method = o->opt.interp_method;
    interp_t it;
    memset(&it, 0, sizeof it);
    it.dim = o->dim;
    it.tol = o->opt.skewness_tol;
    it.maxs = o->opt.stencil_size;
    if (it.maxs < 1) it.maxs = (o->dim == 3) ? 12 : (o->dim == 2 ? 6 : 2);
    if (method == FCFD_INTERP_IDW) it.maxs = 99;
    o->navg_ptr = (int*)malloc(sizeof(int) * (o->nnode + 1));
    int cap = o->nnode * 4 + 16, nent = 0;
    o->navg_idx = (int*)malloc(sizeof(int) * cap);
    o->navg_w = (double*)malloc(sizeof(double) * cap);
    int scap = 64;
    int* ids = (int*)malloc(sizeof(int) * scap);
    V3* pts = (V3*)malloc(sizeof(V3) * scap);
    double* w = (double*)malloc(sizeof(double) * scap);
    for (int n = 0; n < o->nnode; ++n) {
        o->navg_ptr[n] = nent;
        const ivec* nc = &o->node_cells[n];
        if (nc->n > scap) {
            scap = 2 * nc->n;
            ids = (int*)realloc(ids, sizeof(int) * scap); pts = (V3*)realloc(pts, sizeof(V3) * scap);
            w = (double*)realloc(w, sizeof(double) * scap);
        }
        int ns = 0;                                                /* std::set<int>: ascending, unique */
        for (int k = 0; k < nc->n; ++k) ids[ns++] = nc->a[k];
        qsort(ids, ns, sizeof(int), cmp_int);
        int u = 0;
        for (int k = 0; k < ns; ++k) if (u == 0 || ids[k] != ids[u - 1]) ids[u++] = ids[k];
        ns = u;
        for (int k = 0; k < ns; ++k) pts[k] = o->cc[ids[k]];
        V3 point = o->node[n];
        int nw = 0, kind = 0;
        if (method == FCFD_INTERP_WTLI) {
            nw = wtli_weights(&it, ns, point, ids, pts, w, &kind);
        } else if (method == FCFD_INTERP_IDW) {
            double sum = 0.;
            for (int s = 0; s < ns; ++s) { w[s] = vmag(vsub(point, pts[s])); sum += w[s]; }
            for (int s = 0; s < ns; ++s) w[s] /= sum;
            nw = ns;
        } else {
            for (int s = 0; s < ns; ++s) w[s] = 1. / (double)ns;
            nw = ns;
        }
        map_insert_row(&o->navg_idx, &o->navg_w, &cap, &nent, ids, w, nw);
    }
    o->navg_ptr[o->nnode] = nent;
    free(ids); free(pts); free(w);
    return 0;
}

static int face_interpolation_weights(orc_ctx* o) {
    int method = o->opt.interp_method;
    int maxs = o->opt.stencil_size;
    if (maxs < 1) maxs = (o->dim == 3) ? 12 : (o->dim == 2 ? 6 : 2);
    if (method == FCFD_INTERP_IDW) maxs = 99;
    interp_t it;
    memset(&it, 0, sizeof it);
    it.dim = o->dim; it.tol = o->opt.skewness_tol; it.maxs = maxs;
    o->avg_ptr = (int*)malloc(sizeof(int) * (o->nface + 1));
    int cap = o->nface * 2 + 16, nent = 0;
    o->avg_idx = (int*)malloc(sizeof(int) * cap);
    o->avg_w = (double*)malloc(sizeof(double) * cap);
    int* sidx = (int*)malloc(sizeof(int) * 4096);
    V3* spts = (V3*)malloc(sizeof(V3) * 4096);
    int* setbuf = (int*)malloc(sizeof(int) * 4096);
    for (int f = 0; f < o->nface; ++f) {
        o->avg_ptr[f] = nent;
        V3 point = o->fc[f];
        int c = o->fparent[f], n = o->fneigh[f];
        int ns = 0;
        sidx[ns] = c; spts[ns++] = o->cc[c];
        sidx[ns] = n; spts[ns++] = o->cc[n];
        if (ns < maxs) {                                               /* :76-93 */
            int nset = 0;
            for (int k = 0; k < o->cnbr[c].n; ++k) setbuf[nset++] = o->cnbr[c].a[k];
            if (o->fbc[f] == -1) for (int k = 0; k < o->cnbr[n].n; ++k) setbuf[nset++] = o->cnbr[n].a[k];
            qsort(setbuf, nset, sizeof(int), cmp_int);
            int prev = -999999;
            for (int k = 0; k < nset; ++k) {
                int x = setbuf[k];
                if (x == prev) continue;
                prev = x;
                if (x == c || x == n) continue;
                sidx[ns] = x; spts[ns++] = o->cc[x];
            }
        }
        double wts[4096];
        int nw = 0;
        if (method == FCFD_INTERP_WTLI) {
            /* calculate_weights(is_internal = true): sort keeping parent and
             * neighbour first, then tetra / tri / line / point fits */
            int kind = 0;
            nw = wtli_weights(&it, ns, point, sidx, spts, wts, &kind);
        } else if (method == FCFD_INTERP_IDW) {                        /* :109-118 */
            double sum = 0.;
            for (int s = 0; s < ns; ++s) { wts[s] = vmag(vsub(o->fc[f], spts[s])); sum += wts[s]; }
            for (int s = 0; s < ns; ++s) wts[s] /= sum;
            nw = ns;
        } else {                                                       /* :120-124 SIMPLE */
            wts[0] = 0.5; wts[1] = 0.5; nw = 2;
        }
        /* std::map insert order -> ascending keys */
        if (nent + nw > cap) {
            cap = 2 * cap + nw;
            o->avg_idx = (int*)realloc(o->avg_idx, sizeof(int) * cap);
            o->avg_w = (double*)realloc(o->avg_w, sizeof(double) * cap);
        }
        int base = nent;
        for (int s = 0; s < nw; ++s) {
            int key = sidx[s], dup = 0;
            for (int q = base; q < nent; ++q) if (o->avg_idx[q] == key) { dup = 1; break; }
            if (dup) continue;
            int q = nent;
            while (q > base && o->avg_idx[q - 1] > key) { o->avg_idx[q] = o->avg_idx[q - 1]; o->avg_w[q] = o->avg_w[q - 1]; --q; }
            o->avg_idx[q] = key; o->avg_w[q] = wts[s];
            nent++;
        }
    }
    o->avg_ptr[o->nface] = nent;
    free(sidx); free(spts); free(setbuf);
    return 0;
}

/* ------------------------------------------------------------------ */
/* gradient maps (src/gradient_maps.cc:39-78)                          */
/* ------------------------------------------------------------------ */
static void lsqr_grad_map(orc_ctx* o, int ci, vmap* gm) {      /* src/lsqr_grad_map.cc:32-99 */
    static int st[MAPMAX]; static V3 dist[MAPMAX]; static double w[MAPMAX];
    int ns = 0;
    for (int i = 0; i < o->cnbr[ci].n; ++i) { int nb = o->cnbr[ci].a[i]; if (nb != ci && ns < MAPMAX) st[ns++] = nb; }
    double a = 0., b = 0., c = 0., d = 0., e = 0., f = 0.;
    for (int k = 0; k < ns; ++k) {
        dist[k] = vsub(o->cc[st[k]], o->cc[ci]);
        w[k] = vdot(dist[k], dist[k]);
        a += w[k] * dist[k].c[0] * dist[k].c[0];
        b += w[k] * dist[k].c[0] * dist[k].c[1];
        c += w[k] * dist[k].c[0] * dist[k].c[2];
        d += w[k] * dist[k].c[1] * dist[k].c[1];
        e += w[k] * dist[k].c[1] * dist[k].c[2];
        f += w[k] * dist[k].c[2] * dist[k].c[2];
    }
    for (int cf = 0; cf < o->cfaces[ci].n; ++cf) {                      /* :66-81 */
        int fi = o->cfaces[ci].a[cf];
        if (o->fsym[fi] && ns < MAPMAX) {
            st[ns] = ci;
            V3 point = vadd(o->cc[ci], vsc(vdot(vsub(o->fc[fi], o->cc[ci]), o->fn[fi]), o->fn[fi]));
            dist[ns] = vsub(point, o->cc[ci]);
            int k = ns;
            w[k] = vdot(dist[k], dist[k]);
            a += w[k] * dist[k].c[0] * dist[k].c[0];
            b += w[k] * dist[k].c[0] * dist[k].c[1];
            c += w[k] * dist[k].c[0] * dist[k].c[2];
            d += w[k] * dist[k].c[1] * dist[k].c[1];
            e += w[k] * dist[k].c[1] * dist[k].c[2];
            f += w[k] * dist[k].c[2] * dist[k].c[2];
            ns++;
        }
    }
    double det = a * d * f - a * e * e - b * b * f - c * c * d + 2. * b * c * e;
    gm->n = 0;
    for (int k = 0; k < ns; ++k) {                                      /* :87-96 */
        V3 wt;
        wt.c[0] = w[k] * (dist[k].c[0] * (d * f - e * e) + dist[k].c[1] * (c * e - b * f) + dist[k].c[2] * (b * e - c * d));
        wt.c[1] = w[k] * (dist[k].c[0] * (c * e - b * f) + dist[k].c[1] * (a * f - c * c) + dist[k].c[2] * (b * c - a * e));
        wt.c[2] = w[k] * (dist[k].c[0] * (b * e - c * d) + dist[k].c[1] * (b * c - a * e) + dist[k].c[2] * (a * d - b * b));
        wt = vdiv(wt, det);
        vm_insert(gm, st[k], wt);
        int s = vm_find(gm, ci);
        if (s < 0) vm_insert(gm, ci, vmul(wt, -1.));
        else gm->val[s] = vsub(gm->val[s], wt);
    }
}

static int other_cell(const orc_ctx* o, int f, int c) { int x = o->fneigh[f]; if (x == c) x = o->fparent[f]; return x; }

static void curvilinear_grad_map(orc_ctx* o, int c, vmap* gm) { /* src/curvilinear_grad_map.cc:28-228 */
    int face_pairs[6] = {-1, -1, -1, -1, -1, -1};
    int cell_pairs[6] = {-1, -1, -1, -1, -1, -1};
    V3 Jac[3] = {vzero(), vzero(), vzero()};
    int nf = o->cfaces[c].n;
    if (o->cnodes[c].n == 8) {
        /* D3: greedy pairing over the remaining face slots in ascending order */
        int rem[8], nrem = nf;
        for (int i = 0; i < nf; ++i) rem[i] = i;
        int counter = 0;
        while (nrem > 0 && counter < 3) {
            int s1 = rem[0];
            double min_product = 1.e20;
            int f1 = o->cfaces[c].a[s1];
            V3 my = o->fn[f1];
            if (o->fparent[f1] != c) my = vmul(my, -1.);
            face_pairs[2 * counter + 1] = s1;
            for (int k = 1; k < nrem; ++k) {
                int f2 = o->cfaces[c].a[rem[k]];
                V3 other = o->fn[f2];
                if (o->fparent[f2] != c) other = vmul(other, -1.);
                double prod = vdot(my, other);
                if (prod < min_product) { min_product = prod; face_pairs[2 * counter] = rem[k]; }
            }
            int nr = 0;
            for (int k = 0; k < nrem; ++k) if (rem[k] != face_pairs[2 * counter] && rem[k] != face_pairs[2 * counter + 1]) rem[nr++] = rem[k];
            nrem = nr;
            counter++;
        }
        for (int i = 0; i < 3; ++i) {                                    /* :81-91 */
            int cp = other_cell(o, o->cfaces[c].a[face_pairs[2 * i + 1]], c);
            cell_pairs[2 * i + 1] = cp;
            Jac[i] = o->cc[cp];
            int cm = other_cell(o, o->cfaces[c].a[face_pairs[2 * i]], c);
            cell_pairs[2 * i] = cm;
            Jac[i] = vsub(Jac[i], o->cc[cm]);
        }
    } else {                                                             /* :93-155 */
        int st[8], ns = nf;
        for (int i = 0; i < nf; ++i) st[i] = o->cfaces[c].a[i];
        qsort(st, ns, sizeof(int), cmp_int);
        double max_det = 0.;
        int cpt[6];
        V3 Jt[3];
        for (int i1 = 0; i1 < ns; ++i1)
            for (int i2 = i1; i2 < ns; ++i2) {
                cpt[1] = other_cell(o, st[i2], c); Jt[0] = o->cc[cpt[1]];
                cpt[0] = other_cell(o, st[i1], c); Jt[0] = vsub(Jt[0], o->cc[cpt[0]]);
                for (int i3 = i1; i3 < ns; ++i3)
                    for (int i4 = i3; i4 < ns; ++i4) {
                        cpt[3] = other_cell(o, st[i4], c); Jt[1] = o->cc[cpt[3]];
                        cpt[2] = other_cell(o, st[i3], c); Jt[1] = vsub(Jt[1], o->cc[cpt[2]]);
                        for (int i5 = i2; i5 < ns; ++i5)
                            for (int i6 = i5; i6 < ns; ++i6) {
                                cpt[5] = other_cell(o, st[i6], c); Jt[2] = o->cc[cpt[5]];
                                cpt[4] = other_cell(o, st[i5], c); Jt[2] = vsub(Jt[2], o->cc[cpt[4]]);
                                double det = vdot(Jt[0], vcross(Jt[1], Jt[2]));
                                if (det > max_det) {
                                    max_det = det;
                                    for (int q = 0; q < 6; ++q) cell_pairs[q] = cpt[q];
                                    Jac[0] = Jt[0]; Jac[1] = Jt[1]; Jac[2] = Jt[2];
                                }
                            }
                    }
            }
    }
    double det = vdot(Jac[0], vcross(Jac[1], Jac[2]));                  /* :160-172 */
    V3 JiT[3];
    JiT[0].c[0] = (Jac[1].c[1] * Jac[2].c[2] - Jac[1].c[2] * Jac[2].c[1]) / det;
    JiT[1].c[0] = (Jac[0].c[2] * Jac[2].c[1] - Jac[0].c[1] * Jac[2].c[2]) / det;
    JiT[2].c[0] = (Jac[0].c[1] * Jac[1].c[2] - Jac[0].c[2] * Jac[1].c[1]) / det;
    JiT[0].c[1] = (Jac[1].c[2] * Jac[2].c[0] - Jac[1].c[0] * Jac[2].c[2]) / det;
    JiT[1].c[1] = (Jac[0].c[0] * Jac[2].c[2] - Jac[0].c[2] * Jac[2].c[0]) / det;
    JiT[2].c[1] = (Jac[0].c[2] * Jac[1].c[0] - Jac[0].c[0] * Jac[1].c[2]) / det;
    JiT[0].c[2] = (Jac[1].c[0] * Jac[2].c[1] - Jac[1].c[1] * Jac[2].c[0]) / det;
    JiT[1].c[2] = (Jac[0].c[1] * Jac[2].c[0] - Jac[0].c[0] * Jac[2].c[1]) / det;
    JiT[2].c[2] = (Jac[0].c[0] * Jac[1].c[1] - Jac[0].c[1] * Jac[1].c[0]) / det;
    gm->n = 0;                                                           /* :203-220 */
    vm_insert(gm, cell_pairs[0], vmul(JiT[0], -1.));
    for (int q = 1; q < 6; ++q) {
        int dir = q / 2;
        int plus = q & 1;
        int s = vm_find(gm, cell_pairs[q]);
        if (s < 0) vm_insert(gm, cell_pairs[q], plus ? JiT[dir] : vsc(-1., JiT[dir]));
        else gm->val[s] = plus ? vadd(gm->val[s], JiT[dir]) : vsub(gm->val[s], JiT[dir]);
    }
}

static int gradient_maps(orc_ctx* o) {
    o->g_ptr = (int*)malloc(sizeof(int) * (o->ncell + 1));
    int cap = o->ncell * 8 + 64, n = 0;
    o->g_idx = (int*)malloc(sizeof(int) * cap);
    o->g_w = (V3*)malloc(sizeof(V3) * cap);
    vmap* gm = (vmap*)malloc(sizeof(vmap));
    for (int c = 0; c < o->ncell; ++c) {
        o->g_ptr[c] = n;
        int nn = o->cnodes[c].n;
        int meth = (nn == 8) ? o->opt.grad_hex : (nn == 6) ? o->opt.grad_prism : o->opt.grad_other;
        gm->n = 0;
        if (meth == FCFD_GRAD_CURVILINEAR) curvilinear_grad_map(o, c, gm);
        else if (meth == FCFD_GRAD_LSQR) lsqr_grad_map(o, c, gm);
        if (n + gm->n > cap) {
            cap = 2 * cap + gm->n;
            o->g_idx = (int*)realloc(o->g_idx, sizeof(int) * cap);
            o->g_w = (V3*)realloc(o->g_w, sizeof(V3) * cap);
        }
        for (int k = 0; k < gm->n; ++k) { o->g_idx[n] = gm->key[k]; o->g_w[n] = gm->val[k]; n++; }
    }
    o->g_ptr[o->ncell] = n;
    free(gm);
    return 0;
}

/* ------------------------------------------------------------------ */
/* public set-up                                                      */
/* ------------------------------------------------------------------ */
orc_ctx* orc_create(const fcfd_mesh* mesh, const fcfd_options* opt) {
    if (!mesh || !opt || mesh->n_cells <= 0 || mesh->n_bc > FCFD_MAX_BC || opt->n_bc != mesh->n_bc) {
        seterr(NULL, FCFD_EINVAL, "invalid mesh/options");
        return NULL;
    }
    orc_ctx* o = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    o->opt = *opt;
    o->dim = mesh->dimension;
    o->R = 8314.47215 / opt->Mw;
    o->Cp = opt->gamma * o->R / (opt->gamma - 1.);
    int rc = create_nodes_cells(o, mesh);
    if (!rc) rc = create_faces(o);
    if (rc) { snprintf(g_err, sizeof g_err, "%s", o->msg); orc_destroy(o); return NULL; }
    areas_volumes(o);
    create_boundary_ghosts(o);
    set_bcs_grid(o);
    set_length_scales(o);
    rc = face_interpolation_weights(o);
    if (!rc) rc = node_interpolation_weights(o);
    if (!rc) rc = gradient_maps(o);
    if (rc) { snprintf(g_err, sizeof g_err, "%s", o->msg); orc_destroy(o); return NULL; }
    int nt = o->ntot;
    o->p = (double*)calloc(nt, sizeof(double)); o->u = (double*)calloc(nt, sizeof(double));
    o->v = (double*)calloc(nt, sizeof(double)); o->w = (double*)calloc(nt, sizeof(double));
    o->T = (double*)calloc(nt, sizeof(double));
    o->gp = (V3*)calloc(nt, sizeof(V3)); o->gu = (V3*)calloc(nt, sizeof(V3)); o->gv = (V3*)calloc(nt, sizeof(V3));
    o->gw = (V3*)calloc(nt, sizeof(V3)); o->gT = (V3*)calloc(nt, sizeof(V3));
    o->nv = (opt->turbulence == FCFD_TURB_SA) ? 6 : 5;
    o->nut = (double*)calloc(nt, sizeof(double));
    o->gn = (V3*)calloc(nt, sizeof(V3));
    o->phi = (double*)malloc(sizeof(double) * nt * NVM);
    for (int i = 0; i < nt * NVM; ++i) o->phi[i] = (opt->order == FCFD_ORDER_FIRST) ? 0. : 1.;   /* ns.cc:282-287 */
    o->upd = (double*)calloc((size_t)o->ncell * NVM, sizeof(double));
    o->Rs = (double*)calloc((size_t)o->ncell * NVM, sizeof(double));
    o->Rabs = (double*)calloc((size_t)o->ncell * NVM, sizeof(double));
    o->fC = (double*)calloc((size_t)o->nface * NVM, sizeof(double));
    o->fD = (double*)calloc((size_t)o->nface * NVM, sizeof(double));
    o->loads = (double*)calloc((size_t)(o->nbc + 1) * 6, sizeof(double));
    o->wL = (double*)calloc((size_t)o->nface, sizeof(double));
    o->kw = opt->turbulence == FCFD_TURB_KW;
    o->kk = (double*)calloc(nt, sizeof(double)); o->om = (double*)calloc(nt, sizeof(double));
    o->mut = (double*)calloc(nt, sizeof(double)); o->strain = (double*)calloc(nt, sizeof(double));
    o->gk = (V3*)calloc(nt, sizeof(V3)); o->gom = (V3*)calloc(nt, sizeof(V3));
    int nb1 = o->nbc + 1;
    o->kw_kbc = (double*)calloc(nb1, sizeof(double)); o->kw_obc = (double*)calloc(nb1, sizeof(double));
    o->kw_mbc = (double*)calloc(nb1, sizeof(double));
    o->kw_kfix = (int*)calloc(nb1, sizeof(int)); o->kw_mfix = (int*)calloc(nb1, sizeof(int));
    for (int b = 0; b < o->nbc; ++b) {   /* rans_set_bcs.cc:39-79 */
        const fcfd_bc* bc = &opt->bc[b];
        if (bc->type == FCFD_BC_INLET) { o->kw_kfix[b] = 1; o->kw_mfix[b] = 1; }
        else if (bc->type == FCFD_BC_WALL && bc->kind != FCFD_KIND_SLIP) { o->kw_mfix[b] = 1; o->kw_mbc[b] = 0.; }
    }
    return o;
}

void orc_destroy(orc_ctx* o) {
    if (!o) return;
    for (int c = 0; c < o->ncap && o->cnodes; ++c) { iv_free(&o->cnodes[c]); iv_free(&o->cfaces[c]); if (o->cnbr) iv_free(&o->cnbr[c]); }
    for (int n = 0; n < o->nnode && o->node_cells; ++n) iv_free(&o->node_cells[n]);
    for (int f = 0; f < o->nface && o->fnodes; ++f) iv_free(&o->fnodes[f]);
    for (int b = 0; b < o->nbc && o->bcflag; ++b) free(o->bcflag[b]);
    free(o->bcflag); free(o->cnodes); free(o->cfaces); free(o->cnbr); free(o->node_cells); free(o->node);
    free(o->cc); free(o->vol); free(o->lscale); free(o->wdist); free(o->cbc);
    free(o->fparent); free(o->fneigh); free(o->fbc); free(o->fsym); free(o->face2bc); free(o->fnodes);
    free(o->fc); free(o->fn); free(o->farea); free(o->fwd); free(o->fdiss);
    free(o->bc_area); free(o->bc_total_area);
    free(o->avg_ptr); free(o->avg_idx); free(o->avg_w); free(o->g_ptr); free(o->g_idx); free(o->g_w);
    free(o->p); free(o->u); free(o->v); free(o->w); free(o->T);
    free(o->gp); free(o->gu); free(o->gv); free(o->gw); free(o->gT); free(o->nut); free(o->gn);
    free(o->phi); free(o->upd); free(o->Rs); free(o->Rabs); free(o->fC); free(o->fD); free(o->navg_ptr); free(o->navg_idx); free(o->navg_w); free(o->JL); free(o->JR); free(o->Dg); free(o->loads);
    free(o->wL); free(o->kk); free(o->om); free(o->mut); free(o->strain); free(o->gk); free(o->gom);
    free(o->kw_kbc); free(o->kw_obc); free(o->kw_mbc); free(o->kw_kfix); free(o->kw_mfix);
    free(o);
}

int orc_info(const orc_ctx* o, fcfd_info* info) {
    memset(info, 0, sizeof *info);
    info->n_nodes = o->nnode; info->n_cells = o->ncell; info->n_ghosts = o->ntot - o->ncell;
    info->n_faces = o->nface; info->n_bfaces = o->ntot - o->ncell;
    info->ld_cell = o->ntot; info->ld_face = o->nface;
    info->n_gradmap = o->g_ptr[o->ncell]; info->n_faceavg = o->avg_ptr[o->nface];
    int nn = 0; for (int c = 0; c < o->ntot; ++c) nn += o->cnbr[c].n;
    info->n_nbr = nn;
    int hf = 0; for (int c = 0; c < o->ncell; ++c) hf += o->cfaces[c].n;
    info->n_halffaces = hf;
    info->nv = o->nv;
    info->n_nodeavg = o->navg_ptr[o->nnode];
    return 0;
}

/* face tangent1 (ns_assemble_linear_system.cc:337-341) */
static V3 face_t1(const orc_ctx* o, int f) {
    const ivec* fn = &o->fnodes[f];
    if (fn->n == 4)
        return vnorm(vsub(vadd(o->node[fn->a[0]], o->node[fn->a[1]]), vadd(o->node[fn->a[2]], o->node[fn->a[3]])));
    return vnorm(vsub(vsc(0.5, vadd(o->node[fn->a[0]], o->node[fn->a[1]])), o->fc[f]));
}

int orc_metrics(const orc_ctx* o, fcfd_metrics_out* m) {
    int F = o->nface, N = o->ncell, T = o->ntot;
    for (int f = 0; f < F; ++f) {
        if (m->face_parent) m->face_parent[f] = o->fparent[f];
        if (m->face_neighbor) m->face_neighbor[f] = o->fneigh[f];
        if (m->face_bc) m->face_bc[f] = o->fbc[f];
        for (int k = 0; k < 3; ++k) {
            if (m->face_centroid) m->face_centroid[3 * f + k] = o->fc[f].c[k];
            if (m->face_normal) m->face_normal[3 * f + k] = o->fn[f].c[k];
        }
        if (m->face_tangent1) { V3 t = face_t1(o, f); for (int k = 0; k < 3; ++k) m->face_tangent1[3 * f + k] = t.c[k]; }
        if (m->face_area) m->face_area[f] = o->farea[f];
    }
    if (m->face_node_ptr) {
        int n = 0;
        for (int f = 0; f < F; ++f) { m->face_node_ptr[f] = n; for (int k = 0; k < o->fnodes[f].n; ++k) { if (m->face_nodes) m->face_nodes[n] = o->fnodes[f].a[k]; n++; } }
        m->face_node_ptr[F] = n;
    }
    for (int c = 0; c < T; ++c) {
        for (int k = 0; k < 3; ++k) if (m->cell_centroid) m->cell_centroid[3 * c + k] = o->cc[c].c[k];
        if (m->cell_volume) m->cell_volume[c] = o->vol[c];
        if (m->wall_distance) m->wall_distance[c] = o->wdist[c];
        if (m->cell_perm) m->cell_perm[c] = c;
    }
    for (int c = 0; c < N; ++c) if (m->cell_lengthscale) m->cell_lengthscale[c] = o->lscale[c];
    if (m->cell_face_ptr) {
        int n = 0;
        for (int c = 0; c < N; ++c) { m->cell_face_ptr[c] = n; for (int k = 0; k < o->cfaces[c].n; ++k) { if (m->cell_faces) m->cell_faces[n] = o->cfaces[c].a[k]; n++; } }
        m->cell_face_ptr[N] = n;
    }
    if (m->nbr_ptr) {
        int n = 0;
        for (int c = 0; c < T; ++c) { m->nbr_ptr[c] = n; for (int k = 0; k < o->cnbr[c].n; ++k) { if (m->nbr) m->nbr[n] = o->cnbr[c].a[k]; n++; } }
        m->nbr_ptr[T] = n;
    }
    if (m->grad_ptr) {
        for (int c = 0; c <= N; ++c) m->grad_ptr[c] = o->g_ptr[c];
        for (int k = 0; k < o->g_ptr[N]; ++k) {
            if (m->grad_idx) m->grad_idx[k] = o->g_idx[k];
            if (m->grad_w) for (int q = 0; q < 3; ++q) m->grad_w[3 * k + q] = o->g_w[k].c[q];
        }
    }
    if (m->avg_ptr) {
        for (int f = 0; f <= F; ++f) m->avg_ptr[f] = o->avg_ptr[f];
        for (int k = 0; k < o->avg_ptr[F]; ++k) { if (m->avg_idx) m->avg_idx[k] = o->avg_idx[k]; if (m->avg_w) m->avg_w[k] = o->avg_w[k]; }
    }
    if (m->grid_lengthscale) m->grid_lengthscale[0] = o->glscale;
    for (int n = 0; n < o->nnode; ++n)
        if (m->node_xyz) for (int q = 0; q < 3; ++q) m->node_xyz[3 * n + q] = o->node[n].c[q];
    if (m->node_avg_ptr) {
        for (int n = 0; n <= o->nnode; ++n) m->node_avg_ptr[n] = o->navg_ptr[n];
        for (int k = 0; k < o->navg_ptr[o->nnode]; ++k) {
            if (m->node_avg_idx) m->node_avg_idx[k] = o->navg_idx[k];
            if (m->node_avg_w) m->node_avg_w[k] = o->navg_w[k];
        }
    }
    return 0;
}

/* gelimd on an n x n system (n = 3 or 4), row-major a (n*n), b (n): the
 * probe for the pin tests; returns gelimd's status, x untouched on 1 */
int orc_gelimd_probe(int n, const double* a, const double* b, double* x) {
    double A4[4][4], B4[4];
    if (n < 1 || n > 4) return FCFD_EINVAL;
    for (int i = 0; i < n; ++i) { B4[i] = b[i]; for (int j = 0; j < n; ++j) A4[i][j] = a[n * i + j]; }
    return gelimd(n, A4, B4, x);
}

/* ------------------------------------------------------------------ */
/* state                                                              */
/* ------------------------------------------------------------------ */
static void load_state(orc_ctx* o, const double* W, int n) {
    const int nv = o->nv;
    for (int c = 0; c < n; ++c) {
        o->p[c] = W[nv * c]; o->u[c] = W[nv * c + 1]; o->v[c] = W[nv * c + 2]; o->w[c] = W[nv * c + 3]; o->T[c] = W[nv * c + 4];
        if (nv == 6) o->nut[c] = W[nv * c + 5];
    }
}
int orc_set_state(orc_ctx* o, const double* W) { load_state(o, W, o->ncell); return 0; }
int orc_set_full_state(orc_ctx* o, const double* W) { load_state(o, W, o->ntot); return 0; }
int orc_set_update(orc_ctx* o, const double* dW) {
    for (int c = 0; c < o->ncell; ++c) for (int i = 0; i < o->nv; ++i) o->upd[NVM * c + i] = dW[o->nv * c + i];
    return 0;
}

static V3 cellV(const orc_ctx* o, int c) { return v3(o->u[c], o->v[c], o->w[c]); }

/* ------------------------------------------------------------------ */
/* boundary conditions (src/navier_stokes/ns_apply_bcs.cc:25-258)      */
/* ------------------------------------------------------------------ */
static int apply_bcs(orc_ctx* o, cstate* L, cstate* R, const fstate* F) {
    const fcfd_bc* bc = &o->opt.bc[F->bc];
    double g = o->opt.gamma, Pref = o->opt.Pref, Tref = o->opt.Tref;
    if (bc->type == FCFD_BC_INLET) {
        if (bc->kind == FCFD_KIND_VELOCITY) {                          /* :45-88 */
            R->V = v3(bc->V[0], bc->V[1], bc->V[2]);
            double uNL = vdot(L->V, F->normal);
            double uNR = vdot(R->V, F->normal);
            R->a = L->a + 0.5 * (uNL - uNR) * (g - 1.);
            double MachL = -uNL / L->a;
            if (MachL >= 1. && bc->specified != FCFD_SPEC_STATE) { seterr(o, FCFD_EBC, "inlet became supersonic, thermodynamic state needed"); return FCFD_EBC; }
            if (bc->specified == FCFD_SPEC_STATE) { R->p = bc->p; R->T = bc->T; R->rho = bc->rho; }
            else if (bc->specified == FCFD_SPEC_P) { R->p = bc->p; R->rho = g * (R->p + Pref) / (R->a * R->a); R->T = m_T(o, R->p, R->rho); }
            else if (bc->specified == FCFD_SPEC_T) { R->T = bc->T; R->rho = L->rho * pow((L->T + Tref) / (R->T + Tref), 1. / (1. - g)); R->p = m_p(o, R->rho, R->T); }
            else if (bc->specified == FCFD_SPEC_RHO) { R->rho = bc->rho; R->p = R->a * R->a * R->rho / g - Pref; R->T = m_T(o, R->p, R->rho); }
            else { seterr(o, FCFD_EBC, "inlet needs one of p,T,rho"); return FCFD_EBC; }
            R->p_center = R->p; R->V_center = R->V; R->T_center = R->T;
        } else if (bc->kind == FCFD_KIND_MDOT) {                       /* :90-135 */
            double uNL = vdot(L->V, F->normal);
            double mdotNR = -1. * bc->mdot / o->bc_total_area[F->bc];
            /* :99-104: `!specified != BC_STATE` holds for every specification */
            if (-uNL / L->a >= 1.) { seterr(o, FCFD_EBC, "mass-flow inlet became supersonic"); return FCFD_EBC; }
            if (bc->specified == FCFD_SPEC_STATE) { R->p = bc->p; R->T = bc->T; R->rho = bc->rho; }
            else if (bc->specified == FCFD_SPEC_P) { R->p = bc->p; R->T = L->T; R->rho = m_rho(o, R->p, R->T); }
            else if (bc->specified == FCFD_SPEC_T) { R->T = bc->T; R->p = L->p; R->rho = m_rho(o, R->p, R->T); }
            else if (bc->specified == FCFD_SPEC_RHO) {
                R->rho = bc->rho;
                R->a = 0.5 * (g - 1.) * (uNL + 2. * L->a / (g - 1.) - mdotNR / R->rho);
                R->p = R->a * R->a * R->rho / g - Pref;
                R->T = m_T(o, R->p, R->rho);
            } else { seterr(o, FCFD_EBC, "inlet needs one of p,T,rho"); return FCFD_EBC; }
            R->V = vsc(mdotNR / R->rho, F->normal);
            R->p_center = R->p; R->V_center = R->V; R->T_center = R->T;
        } else if (bc->kind == FCFD_KIND_STAGNATION) {                 /* :137-155 */
            if (vdot(L->V, F->normal) > 0.) R->V = vsc(-1., L->V_center);
            else R->V = L->V_center;
            R->T = bc->T_total - 0.5 * (g - 1.) / (g * o->R) * vdot(R->V, R->V);
            R->p = (bc->p_total + Pref) / (1. + 0.5 * vdot(R->V, R->V) / (o->R * (R->T + Tref))) - Pref;
            R->rho = m_rho(o, R->p, R->T);
            R->p_center = R->p; R->V_center = R->V; R->T_center = R->T;
        }
    } else if (bc->type == FCFD_BC_OUTLET) {                           /* :157-215 */
        double uNL = vdot(L->V, F->normal);
        double MachL = uNL / L->a;
        int supersonic = (MachL > 1.);
        if (bc->specified == FCFD_SPEC_STATE) { seterr(o, FCFD_EBC, "complete thermodynamic state at an outlet"); return FCFD_EBC; }
        else if (bc->specified == FCFD_SPEC_P && !supersonic) {
            R->p = R->p_center = bc->p;
            R->rho = L->rho * pow((R->p + Pref), 1. / g) / pow((L->p + Pref), 1. / g);
            R->a = sqrt(g * (R->p + Pref) / R->rho);
            R->T = m_T(o, R->p, R->rho);
            R->V = L->V;
        } else if (bc->specified == FCFD_SPEC_T && !supersonic) {
            R->T = R->T_center = bc->T;
            R->a = sqrt(g * o->R * (R->T + Tref));
            R->rho = pow(g * (L->p + Pref) / (R->a * R->a * pow(L->rho, g)), 1. - g);
            R->p = m_p(o, R->rho, R->T);
            R->V = L->V;
        } else if (bc->specified == FCFD_SPEC_RHO && !supersonic) {
            R->rho = bc->rho;
            R->p = (L->p + Pref) * pow(R->rho / L->rho, g) - Pref;
            R->T = m_T(o, R->p, R->rho);
            R->a = sqrt(g * (R->p + Pref) / R->rho);
            R->V = L->V;
        } else {
            R->p = L->p; R->T = L->T; R->V = L->V;
            R->rho = m_rho(o, R->p, R->T);      /* D2 */
        }
        if (MachL < 0.) {
            if (bc->kind == FCFD_KIND_DAMP_REVERSE) {
                R->p -= 0.5 * L->rho * uNL * uNL;
                R->rho = L->rho * pow((R->p + Pref), 1. / g) / pow((L->p + Pref), 1. / g);
                R->a = sqrt(g * (R->p + Pref) / R->rho);
                R->T = m_T(o, R->p, R->rho);
            }
            if (bc->kind == FCFD_KIND_NO_REVERSE) R->V = vsc(-1., L->V);
        }
        if (MachL < 1. && bc->kind == FCFD_KIND_FORCE_SUPERSONIC) {
            R->p -= 0.5 * L->rho * pow(L->a - uNL, 2.);
            R->p = dmin(1.e-8, R->p + Pref) - Pref;
            R->rho = L->rho * pow((R->p + Pref), 1. / g) / pow((L->p + Pref), 1. / g);
            R->a = sqrt(g * (R->p + Pref) / R->rho);
            R->T = m_T(o, R->p, R->rho);
        }
    } else if (bc->type == FCFD_BC_WALL) {                             /* :217-244 */
        if (bc->thermal == FCFD_THERMAL_FIXED_T) {
            R->T = R->T_center = bc->T;
            R->p = L->p;
            R->rho = m_rho(o, R->p, R->T);
        } else {
            R->p = L->p; R->rho = L->rho; R->T = L->T; R->T_center = L->T_center;
        }
        if (bc->kind == FCFD_KIND_SLIP) {
            R->V = vsub(L->V, vsc(2. * vdot(L->V, F->normal), F->normal));
            R->V_center = vsub(L->V_center, vsc(2. * vdot(L->V_center, F->normal), F->normal));
        } else {
            R->V = vsc(-1., L->V);
            R->V_center = vsc(-1., L->V_center);
        }
        R->p_center = L->p_center;
    } else if (bc->type == FCFD_BC_SYMMETRY) {                         /* :246-258 */
        R->p = L->p;
        R->V = vsub(L->V, vsc(2. * vdot(L->V, F->normal), F->normal));
        R->T = L->T;
        R->rho = L->rho;
        R->p_center = L->p_center;
        R->T_center = L->T_center;
        R->V_center = vsub(L->V_center, vsc(2. * vdot(L->V_center, F->normal), F->normal));
    }
    /* SA nu~ (build-only): inlet value; zero on a no-slip wall (the ghost
     * mirrors to -nu~ so the face average vanishes); copied otherwise */
    if (bc->type == FCFD_BC_INLET) { R->nut = bc->nut; R->nut_center = bc->nut; }
    else if (bc->type == FCFD_BC_WALL && bc->kind != FCFD_KIND_SLIP) { R->nut = -1. * L->nut; R->nut_center = -1. * L->nut_center; }
    else if (bc->type == FCFD_BC_OUTLET) { R->nut = L->nut; }
    else { R->nut = L->nut; R->nut_center = L->nut_center; }
    return 0;
}

/* face_geom_update (ns_assemble_linear_system.cc:332-351) */
static void face_geom_update(const orc_ctx* o, fstate* F, int f) {
    F->index = f;
    F->parent = o->fparent[f];
    F->neighbor = o->fneigh[f];
    F->normal = o->fn[f];
    F->tangent1 = face_t1(o, f);
    F->tangent2 = vnorm(vcross(F->normal, F->tangent1));
    F->area = o->farea[f];
    F->bc = o->fbc[f];
    F->left2right = vsub(o->cc[F->neighbor], o->cc[F->parent]);
}

static void finish_state(const orc_ctx* o, cstate* S, const fstate* F) {   /* :286-290, :323-327 */
    S->a = m_a(o, S->p, S->T);
    S->H = S->a * S->a / (o->opt.gamma - 1.) + 0.5 * vdot(S->V, S->V);
    S->Vn = v3(vdot(S->V, F->normal), vdot(S->V, F->tangent1), vdot(S->V, F->tangent2));
}

/* left_state_update / interior part of right_state_update (:269-294, :305-320) */
static void recon_state(const orc_ctx* o, cstate* S, const fstate* F, int c) {
    V3 c2f = vsub(vsc(1., o->fc[F->index]), o->cc[c]);    /* order_factor == 1 in the face loop */
    const double* ph = &o->phi[NVM * c];
    S->p_center = o->p[c];
    S->p = S->p_center + ph[0] * vdot(c2f, o->gp[c]);
    V3 dV = v3(ph[1] * vdot(c2f, o->gu[c]), ph[2] * vdot(c2f, o->gv[c]), ph[3] * vdot(c2f, o->gw[c]));
    S->V_center = cellV(o, c);
    S->V = vadd(S->V_center, dV);
    S->T_center = o->T[c];
    S->T = S->T_center + ph[4] * vdot(c2f, o->gT[c]);
    S->rho = m_rho(o, S->p, S->T);
    S->nut_center = o->nut[c];
    S->nut = (o->nv == 6) ? S->nut_center + ph[5] * vdot(c2f, o->gn[c]) : 0.;
    for (int i = 0; i < NVM; ++i) S->update[i] = (c < o->ncell) ? o->upd[NVM * c + i] : 0.;
    S->volume = o->vol[c];
}

static int right_state_update(orc_ctx* o, cstate* L, cstate* R, const fstate* F) {   /* :296-330 */
    if (F->bc >= 0) {
        for (int i = 0; i < NVM; ++i) R->update[i] = 0.;
        int g = F->neighbor;
        R->p_center = o->p[g]; R->T_center = o->T[g]; R->V_center = cellV(o, g);
        R->nut_center = o->nut[g];
        int rc = apply_bcs(o, L, R, F);
        if (rc) return rc;
    } else {
        recon_state(o, R, F, F->neighbor);
    }
    finish_state(o, R, F);
    return 0;
}

/* SA face quantities: face gradient of nu~ (face average + the l2r
 * correction of face_state_update), face nu~ and rho (averages of the two
 * face states) and mu_t(face).  with_grad = 0 keeps the gradient. */
static void sa_face_update(const orc_ctx* o, const cstate* L, const cstate* R, fstate* F, int with_grad) {
    if (with_grad) {
        if (F->bc >= 0) F->gradn = o->gn[F->parent];
        else {
            V3 sn = vzero();
            for (int k = o->avg_ptr[F->index]; k < o->avg_ptr[F->index + 1]; ++k) sn = vadd(sn, vsc(o->avg_w[k], o->gn[o->avg_idx[k]]));
            F->gradn = sn;
        }
        V3 l2rn = vnorm(F->left2right);
        double l2rmag = vmag(F->left2right);
        F->gradn = vsub(F->gradn, vsc(vdot(F->gradn, l2rn), l2rn));
        F->gradn = vadd(F->gradn, vsc((R->nut_center - L->nut_center) / l2rmag, l2rn));
    }
    F->nut = 0.5 * (L->nut + R->nut);
    F->rho = 0.5 * (L->rho + R->rho);
    F->mu_t = sa_mut(F->rho, F->nut, F->mu);
}

/* Variable::face_calculate (variable.h:134-147) of a RANS scalar: the BC
 * value where it is fixed on the face's BC, else the face average */
static double kw_face_value(const orc_ctx* o, const double* q, const int* fixed, const double* bcv, int f) {
    int b = o->fbc[f];
    if (b >= 0 && fixed[b]) return bcv[b];
    double t = 0.;
    for (int k = o->avg_ptr[f]; k < o->avg_ptr[f + 1]; ++k) t += o->avg_w[k] * q[o->avg_idx[k]];
    return t;
}

/* face_state_update (:353-415) */
static void face_state_update(const orc_ctx* o, const cstate* L, const cstate* R, fstate* F) {
    if (F->bc >= 0) {
        F->gradu = o->gu[F->parent]; F->gradv = o->gv[F->parent]; F->gradw = o->gw[F->parent]; F->gradT = o->gT[F->parent];
    } else {   /* Variable<Vec3D>::face_calculate (variable.h:134-147) */
        V3 su = vzero(), sv = vzero(), sw = vzero(), sT = vzero();
        for (int k = o->avg_ptr[F->index]; k < o->avg_ptr[F->index + 1]; ++k) {
            int c = o->avg_idx[k];
            double wt = o->avg_w[k];
            su = vadd(su, vsc(wt, o->gu[c])); sv = vadd(sv, vsc(wt, o->gv[c]));
            sw = vadd(sw, vsc(wt, o->gw[c])); sT = vadd(sT, vsc(wt, o->gT[c]));
        }
        F->gradu = su; F->gradv = sv; F->gradw = sw; F->gradT = sT;
    }
    V3 l2rn = vnorm(F->left2right);
    double l2rmag = vmag(F->left2right);
    F->gradu = vsub(F->gradu, vsc(vdot(F->gradu, l2rn), l2rn));
    F->gradu = vadd(F->gradu, vsc((R->V_center.c[0] - L->V_center.c[0]) / l2rmag, l2rn));
    F->gradv = vsub(F->gradv, vsc(vdot(F->gradv, l2rn), l2rn));
    F->gradv = vadd(F->gradv, vsc((R->V_center.c[1] - L->V_center.c[1]) / l2rmag, l2rn));
    F->gradw = vsub(F->gradw, vsc(vdot(F->gradw, l2rn), l2rn));
    F->gradw = vadd(F->gradw, vsc((R->V_center.c[2] - L->V_center.c[2]) / l2rmag, l2rn));
    F->gradT = vsub(F->gradT, vsc(vdot(F->gradT, l2rn), l2rn));
    F->gradT = vadd(F->gradT, vsc((R->T_center - L->T_center) / l2rmag, l2rn));
    F->V = vsc(0.5, vadd(L->V, R->V));
    F->T = 0.5 * (L->T + R->T);
    F->mu = m_visc(o, F->T);
    F->lambda = m_cond(o, F->T);
    F->gradn = vzero(); F->nut = 0.; F->rho = 0.; F->mu_t = 0.;
    if (o->nv == 6) sa_face_update(o, L, R, F, 1);
    if (o->kw) F->mu_t = kw_face_value(o, o->mut, o->kw_mfix, o->kw_mbc, F->index);   /* rans[gid].mu_t.face (ns_diffusive_face_flux.cc:37) */
}

/* face_state_adjust (:460-500), including the case-1 fall-through */
static void face_state_adjust(const orc_ctx* o, const cstate* L, const cstate* R, fstate* F, int var) {
    V3 l2rn = vnorm(F->left2right);
    double l2rmag = vmag(F->left2right);
    switch (var) {
    case 0: break;
    case 1:
        F->V.c[0] = 0.5 * (L->V.c[0] + R->V.c[0]);
        F->gradu = vsub(F->gradu, vsc(vdot(F->gradu, l2rn), l2rn));
        F->gradu = vadd(F->gradu, vsc((R->V_center.c[0] - L->V_center.c[0]) / l2rmag, l2rn));
        /* fall through */
    case 2:
        F->V.c[1] = 0.5 * (L->V.c[1] + R->V.c[1]);
        F->gradv = vsub(F->gradv, vsc(vdot(F->gradv, l2rn), l2rn));
        F->gradv = vadd(F->gradv, vsc((R->V_center.c[1] - L->V_center.c[1]) / l2rmag, l2rn));
        break;
    case 3:
        F->V.c[2] = 0.5 * (L->V.c[2] + R->V.c[2]);
        F->gradw = vsub(F->gradw, vsc(vdot(F->gradw, l2rn), l2rn));
        F->gradw = vadd(F->gradw, vsc((R->V_center.c[2] - L->V_center.c[2]) / l2rmag, l2rn));
        break;
    case 4:
        F->T = 0.5 * (L->T + R->T);
        F->gradT = vsub(F->gradT, vsc(vdot(F->gradT, l2rn), l2rn));
        F->gradT = vadd(F->gradT, vsc((R->T_center - L->T_center) / l2rmag, l2rn));
        F->mu = m_visc(o, F->T);
        F->lambda = m_cond(o, F->T);
        break;
    }
    if (o->nv == 6) {   /* SA face quantities follow the perturbed states */
        if (var == 5) {
            V3 l2rn2 = vnorm(F->left2right);
            double l2rmag2 = vmag(F->left2right);
            F->gradn = vsub(F->gradn, vsc(vdot(F->gradn, l2rn2), l2rn2));
            F->gradn = vadd(F->gradn, vsc((R->nut_center - L->nut_center) / l2rmag2, l2rn2));
        }
        sa_face_update(o, L, R, F, 0);
    }
}

/* state_perturb (:417-458) */
static void state_perturb(const orc_ctx* o, cstate* S, const fstate* F, int var, double eps) {
    switch (var) {
    case 0: S->p += eps; S->rho = m_rho(o, S->p, S->T); S->a = m_a(o, S->p, S->T); break;
    case 1: case 2: case 3:
        S->V.c[var - 1] += eps; S->V_center.c[var - 1] += eps;
        S->Vn = v3(vdot(S->V, F->normal), vdot(S->V, F->tangent1), vdot(S->V, F->tangent2));
        break;
    case 4: S->T += eps; S->T_center += eps; S->rho = m_rho(o, S->p, S->T); S->a = m_a(o, S->p, S->T); break;
    case 5: S->nut += eps; S->nut_center += eps; break;
    }
    S->H = S->a * S->a / (o->opt.gamma - 1.) + 0.5 * vdot(S->V, S->V);
}

/* ------------------------------------------------------------------ */
/* convective fluxes                                                  */
/* ------------------------------------------------------------------ */
static void roe_flux(const cstate* L, const cstate* R, double fn[5], double Gamma) {   /* ns_roe.cc:25-98 */
    double rho = sqrt(R->rho / L->rho);
    double u = (L->Vn.c[0] + rho * R->Vn.c[0]) / (1. + rho);
    double v = (L->Vn.c[1] + rho * R->Vn.c[1]) / (1. + rho);
    double w = (L->Vn.c[2] + rho * R->Vn.c[2]) / (1. + rho);
    double H = (L->H + rho * R->H) / (1. + rho);
    double a = sqrt((Gamma - 1.) * (H - 0.5 * (u * u + v * v + w * w)));
    rho *= L->rho;
    double Du = R->Vn.c[0] - L->Vn.c[0];
    double Dp = R->p - L->p;
    double deltaV, lambda, Dlambda, mdot, product;
    if (u >= 0.) {
        deltaV = 0.5 * (Dp - rho * a * Du) / (a * a);
        lambda = u - a;
        Dlambda = 2. * (dmin(a, dmax(0., 2. * (L->a - R->a + Du))));
        if (lambda <= (-0.5 * Dlambda)) { ; }
        else if (lambda < (0.5 * Dlambda)) lambda = -0.5 * (lambda - 0.5 * Dlambda) * (lambda - 0.5 * Dlambda) / Dlambda;
        else lambda = 0.;
        mdot = L->rho * L->Vn.c[0];
        product = lambda * deltaV;
        fn[0] = mdot + product;
        fn[1] = mdot * L->Vn.c[0] + L->p + product * (u - a);
        fn[2] = mdot * L->Vn.c[1] + product * v;
        fn[3] = mdot * L->Vn.c[2] + product * w;
        fn[4] = mdot * L->H + product * (H - u * a);
    } else {
        deltaV = 0.5 * (Dp + rho * a * Du) / (a * a);
        lambda = u + a;
        Dlambda = 2. * (dmin(a, dmax(0., 2. * (R->a - L->a + Du))));
        if (lambda >= (0.5 * Dlambda)) { ; }
        else if (lambda > (-0.5 * Dlambda)) lambda = 0.5 * (lambda + 0.5 * Dlambda) * (lambda + 0.5 * Dlambda) / Dlambda;
        else lambda = 0.;
        mdot = R->rho * R->Vn.c[0];
        product = lambda * deltaV;
        fn[0] = mdot - product;
        fn[1] = mdot * R->Vn.c[0] + R->p - product * (u + a);
        fn[2] = mdot * R->Vn.c[1] - product * v;
        fn[3] = mdot * R->Vn.c[2] - product * w;
        fn[4] = mdot * R->H - product * (H + u * a);
    }
}

/* AUSM+-up (ns_ausm_plus_up.cc:35-152) */
static const double ausm_beta = 0.125;
static double msp2(double M) { return 0.25 * (M + 1.) * (M + 1.); }
static double msm2(double M) { return -0.25 * (M - 1.) * (M - 1.); }
static double msp4(double M) { return (fabs(M) >= 1.) ? 0.5 * (M + fabs(M)) : msp2(M) * (1. - 16. * ausm_beta * msm2(M)); }
static double msm4(double M) { return (fabs(M) >= 1.) ? 0.5 * (M - fabs(M)) : msm2(M) * (1. + 16. * ausm_beta * msp2(M)); }
static double psp5(double M, double alpha) { return (fabs(M) >= 1.) ? 0.5 * (M + fabs(M)) / M : msp2(M) * ((2. - M) - 16. * alpha * M * msm2(M)); }
static double psm5(double M, double alpha) { return (fabs(M) >= 1.) ? 0.5 * (M - fabs(M)) / M : msm2(M) * ((-2. - M) + 16. * alpha * M * msp2(M)); }

static void ausm_flux(const cstate* L, const cstate* R, double fn[5], double Pref, double Minf, double* alpha_io) {
    double Kp = 0.25, Ku = 0.75, sigma = 1.;
    double aL = L->a, aR = R->a;
    double aLh = aL * aL / dmax(aL, L->Vn.c[0]);
    double aRh = aR * aR / dmax(aR, -1. * R->Vn.c[0]);
    double a = dmin(aLh, aRh);
    double rho = 0.5 * (L->rho + R->rho);
    double ML = L->Vn.c[0] / a, MR = R->Vn.c[0] / a;
    double Mbar2 = 0.5 * (ML * ML + MR * MR);
    double Mref = dmax(Minf, sqrt(Mbar2));
    Mref = dmin(Mref, 1.);
    double fa;
    if (Mbar2 >= 1.) fa = 1.;
    else { double Mo = sqrt(dmin(1., dmax(Mbar2, Mref * Mref))); fa = Mo * (2. - Mo); }
    /* the global `alpha` (ns_ausm_plus_up.cc:26) still holds the previous
     * call's value while msp4/msm4 run; p_split_5 uses the updated one */
    double M = msp4(ML) + msm4(MR) - Kp / fa * dmax(1. - sigma * Mbar2, 0.) * (R->p - L->p) / (rho * a * a);
    double mdot = (M > 0) ? a * M * L->rho : a * M * R->rho;
    double alpha = 3. / 16. * (-4. + 5. * fa * fa);
    *alpha_io = alpha;
    double p = psp5(ML, alpha) * (L->p + Pref) + psm5(MR, alpha) * (R->p + Pref)
             - Ku * psp5(ML, alpha) * psm5(MR, alpha) * (L->rho + R->rho) * fa * a * (R->Vn.c[0] - L->Vn.c[0]);
    p -= Pref;
    fn[0] = mdot;
    if (mdot > 0.) {
        fn[1] = mdot * L->Vn.c[0] + p; fn[2] = mdot * L->Vn.c[1]; fn[3] = mdot * L->Vn.c[2]; fn[4] = mdot * L->H;
    } else {
        fn[1] = mdot * R->Vn.c[0] + p; fn[2] = mdot * R->Vn.c[1]; fn[3] = mdot * R->Vn.c[2]; fn[4] = mdot * R->H;
    }
}

static void vanleer_flux(const cstate* L, const cstate* R, double fn[5], double Gamma, double Pref) {   /* ns_vanleer.cc:25-66 */
    double ML = L->Vn.c[0] / L->a, MR = R->Vn.c[0] / R->a;
    double M = 0.5 * (ML + MR);
    if (M <= -1.) {
        double mdot = R->rho * R->Vn.c[0];
        fn[0] = mdot; fn[1] = mdot * R->Vn.c[0] + R->p; fn[2] = mdot * R->Vn.c[1]; fn[3] = mdot * R->Vn.c[2]; fn[4] = mdot * R->H;
    } else if (M >= 1.) {
        double mdot = L->rho * L->Vn.c[0];
        fn[0] = mdot; fn[1] = mdot * L->Vn.c[0] + L->p; fn[2] = mdot * L->Vn.c[1]; fn[3] = mdot * L->Vn.c[2]; fn[4] = mdot * L->H;
    } else {
        double fplus = 0.25 * L->rho * L->a * (ML + 1.) * (ML + 1.);
        double fminus = -0.25 * R->rho * R->a * (1. - MR) * (1. - MR);
        fn[0] = fplus + fminus;
        double termL = (2. * L->a / Gamma) * (0.5 * (Gamma - 1) * ML + 1.);
        double termR = (2. * R->a / Gamma) * (0.5 * (Gamma - 1) * MR - 1.);
        fn[1] = fplus * termL + fminus * termR - Pref;
        fn[2] = fplus * L->Vn.c[1] + fminus * R->Vn.c[1];
        fn[3] = fplus * L->Vn.c[2] + fminus * R->Vn.c[2];
        termL = 0.5 * Gamma * Gamma / (Gamma * Gamma - 1.) * termL * termL + 0.5 * (L->Vn.c[1] * L->Vn.c[1] + L->Vn.c[2] * L->Vn.c[2]);
        termR = 0.5 * Gamma * Gamma / (Gamma * Gamma - 1.) * termR * termR + 0.5 * (R->Vn.c[1] * R->Vn.c[1] + R->Vn.c[2] * R->Vn.c[2]);
        fn[4] = fplus * termL + fminus * termR;
    }
}

static double sgn(double n) { return (n < 0) ? -1. : (n > 0) ? 1. : 0.; }
static void slau_flux(const cstate* L, const cstate* R, double fn[5], double Pref) {   /* ns_sd_slau.cc:37-94 */
    const double Csd1 = 0.1, Csd2 = 10.;
    double a = 0.5 * (L->a + R->a);
    double Mhat = dmin(1., sqrt(0.5 * (vdot(L->V, L->V) + vdot(R->V, R->V)) / a));
    double chi = (1. - Mhat) * (1. - Mhat);
    double Mp = L->Vn.c[0] / a, Mm = R->Vn.c[0] / a;
    double alpha = 0.;
    double Bp, Bm;
    if (fabs(Mp) < 1) Bp = 0.25 * (2. - Mp) * (Mp + 1.) * (Mp + 1.) + alpha * Mp * (Mp - 1.) * (Mp - 1.);
    else Bp = 0.5 * (1. + sgn(Mp));
    if (fabs(Mm) < 1) Bm = 0.25 * (2. + Mm) * (Mm - 1.) * (Mm - 1.) - alpha * Mm * (Mm - 1.) * (Mm - 1.);
    else Bm = 0.5 * (1. + sgn(-Mm));
    double ave_p = 0.5 * (L->p + R->p) + Pref;
    double delta_p = R->p - L->p;
    double p = ave_p - 0.5 * (Bp - Bm) * delta_p + (1. - chi) * (Bp + Bm - 1.) * ave_p;
    double max_delta_p = fabs(delta_p);
    double gg = -dmax(dmin(Mp, 0.), -1.) * dmin(dmax(Mm, 0.), 1.);
    double Vn = (L->rho * fabs(L->Vn.c[0]) + R->rho * fabs(R->Vn.c[0])) / (L->rho + R->rho);
    double Vnp = (1. - gg) * Vn + gg * fabs(L->Vn.c[0]);
    double Vnm = (1. - gg) * Vn + gg * fabs(R->Vn.c[0]);
    double theta = (Csd2 * fabs(delta_p) / ave_p + Csd1) / (max_delta_p / ave_p + Csd1);
    theta = dmin(1., theta * theta);
    double mdot = 0.5 * (L->rho * (L->Vn.c[0] + Vnp) + R->rho * (R->Vn.c[0] - Vnm) - theta * dmax(0., (1. - Vn / a)) * delta_p / a);
    p -= Pref;
    fn[0] = mdot;
    if (mdot > 0.) { fn[1] = mdot * L->Vn.c[0] + p; fn[2] = mdot * L->Vn.c[1]; fn[3] = mdot * L->Vn.c[2]; fn[4] = mdot * L->H; }
    else { fn[1] = mdot * R->Vn.c[0] + p; fn[2] = mdot * R->Vn.c[1]; fn[3] = mdot * R->Vn.c[2]; fn[4] = mdot * R->H; }
}

/* modified Steger-Warming (ns_stegger_warming.cc:23-145) */
static void sw_flux(const orc_ctx* o, const cstate* L, const cstate* R, double diss, double cwd, double fn[5]) {
    double alpha = 6.;
    double RW = 8314.47215 / o->opt.Mw;
    double gm = o->opt.gamma;
    double wdiss = o->opt.wdiss, blh = o->opt.bl_height;
    double Lm[5][5], Rm[5][5], Jac[5][5], Lam[5], Q[5];
    for (int i = 0; i < 5; ++i) fn[i] = 0.;
    double deltap = fabs(L->p - R->p) / dmin(L->p, R->p);
    double w = 0.5 / (alpha * alpha * deltap * deltap + 1.);
    for (int h = 0; h <= 1; h++) {
        double signal = pow(-1., h);
        double wL = (signal > 0.) ? 1. - w : w;
        double wR = 1. - wL;
        double p = wL * L->p + wR * R->p;
        V3 V = vadd(vsc(wL, L->Vn), vsc(wR, R->Vn));
        double V2 = vmag(V); V2 *= V2;
        double T = wL * L->T + wR * R->T;
        double rho = wL * L->rho + wR * R->rho;
        double Cp = o->Cp;
        double beta = gm * RW / Cp;
        double a = sqrt((1. + beta) * p / rho);
        double a2 = a * a;
        double H = 0.5 * V2 + a2 / (gm - 1.);
        double gs = RW * T + beta * (0.5 * V2 - Cp * T / gm);
        double eta = (beta * V2 - gs) / (beta * a2);
        for (int i = 0; i < 5; ++i) { Lam[i] = 0.; for (int j = 0; j < 5; ++j) { Lm[i][j] = 0.; Rm[i][j] = 0.; } Jac[i][i] = 0.; }
        Lm[0][0] = 1.0 / a2;
        Lm[0][3] = Lm[0][4] = 0.5 / a2;
        Lm[1][0] = V.c[0] / a2; Lm[2][0] = V.c[1] / a2; Lm[3][0] = V.c[2] / a2;
        Lm[1][3] = 0.5 * (V.c[0] + a) / a2;
        Lm[1][4] = 0.5 * (V.c[0] - a) / a2;
        Lm[2][1] = 1.;
        Lm[2][3] = Lm[2][4] = 0.5 * V.c[1] / a2;
        Lm[3][2] = 1.;
        Lm[3][3] = Lm[3][4] = 0.5 * V.c[2] / a2;
        Lm[4][0] = eta; Lm[4][1] = V.c[1]; Lm[4][2] = V.c[2];
        Lm[4][3] = 0.5 * (H + a * V.c[0]) / a2;
        Lm[4][4] = 0.5 * (H - a * V.c[0]) / a2;
        double eps = (cwd > blh) ? wdiss * (a + vmag(V)) : wdiss * diss * (a + fabs(V.c[0]));
        Lam[0] = Lam[1] = Lam[2] = 0.5 * (V.c[0] + signal * sqrt(V.c[0] * V.c[0] + eps * eps));
        Lam[3] = 0.5 * (V.c[0] + a + signal * sqrt((V.c[0] + a) * (V.c[0] + a) + eps * eps));
        Lam[4] = 0.5 * (V.c[0] - a + signal * sqrt((V.c[0] - a) * (V.c[0] - a) + eps * eps));
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Rm[i][j] += Lm[i][j] * Lam[j];   /* mtx_vec_mult into zeroed R */
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Lm[i][j] = Rm[i][j];
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Rm[i][j] = 0.;
        Rm[0][0] = a2 - gs; Rm[0][1] = beta * V.c[0]; Rm[0][2] = beta * V.c[1]; Rm[0][3] = beta * V.c[2]; Rm[0][4] = -beta;
        Rm[1][0] = -V.c[1]; Rm[1][2] = 1.;
        Rm[2][0] = -V.c[2]; Rm[2][3] = 1.;
        Rm[3][0] = gs - V.c[0] * a; Rm[3][1] = a - beta * V.c[0]; Rm[3][2] = -beta * V.c[1]; Rm[3][3] = -beta * V.c[2];
        Rm[4][0] = gs + V.c[0] * a; Rm[4][1] = -a - beta * V.c[0]; Rm[4][2] = -beta * V.c[1]; Rm[4][3] = -beta * V.c[2];
        Rm[3][4] = Rm[4][4] = beta;
        for (int i = 0; i < 5; i++) for (int j = 0; j < 5; j++) { Jac[i][j] = 0.; for (int k = 0; k < 5; k++) Jac[i][j] += Lm[i][k] * Rm[k][j]; }
        const cstate* S = (signal > 0) ? L : R;
        Q[0] = S->rho; Q[1] = S->rho * S->Vn.c[0]; Q[2] = S->rho * S->Vn.c[1]; Q[3] = S->rho * S->Vn.c[2]; Q[4] = S->rho * S->H - S->p;
        for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) fn[i] += Jac[i][j] * Q[j];
    }
}

/* weightL.face, as each flux function leaves it (the segregated RANS
 * convects k and omega with it, rans_terms.cc:79-84): Roe 0.5
 * (ns_roe.cc:95); AUSM+-up and SD-SLAU 1 where the face mass flux is
 * positive, else 0 (ns_ausm_plus_up.cc:93,99, ns_sd_slau.cc:84,90); van Leer
 * 0 for M <= -1, 1 for M >= 1, else 0.5 (ns_vanleer.cc:39,48,62);
 * Steger-Warming w = 0.5/(alpha^2 dp^2 + 1) of its second pass
 * (ns_stegger_warming.cc:38-46).  fn0 is the normal mass flux. */
static double weight_left(const orc_ctx* o, const cstate* L, const cstate* R, double fn0) {
    switch (o->opt.convective_flux) {
    case FCFD_FLUX_ROE: return 0.5;
    case FCFD_FLUX_VAN_LEER: {
        double M = 0.5 * (L->Vn.c[0] / L->a + R->Vn.c[0] / R->a);
        return (M <= -1.) ? 0. : (M >= 1.) ? 1. : 0.5;
    }
    case FCFD_FLUX_SW: {
        double deltap = fabs(L->p - R->p) / dmin(L->p, R->p);
        return 0.5 / (6. * 6. * deltap * deltap + 1.);
    }
    default: return (fn0 > 0.) ? 1. : 0.;
    }
}

/* convective_face_flux (ns_convective_face_flux.cc:32-59); wl (may be NULL)
 * receives weightL */
static void convective_face_flux(orc_ctx* o, const cstate* L, const cstate* R, const fstate* F, double flux[NVM], double* wl) {
    double fn[5];
    switch (o->opt.convective_flux) {
    case FCFD_FLUX_ROE: roe_flux(L, R, fn, o->opt.gamma); break;
    case FCFD_FLUX_VAN_LEER: vanleer_flux(L, R, fn, o->opt.gamma, o->opt.Pref); break;
    case FCFD_FLUX_SD_SLAU: slau_flux(L, R, fn, o->opt.Pref); break;
    case FCFD_FLUX_SW: sw_flux(o, L, R, o->fdiss[F->index], o->fwd[F->index], fn); break;
    default: { double alpha = 0.; ausm_flux(L, R, fn, o->opt.Pref, o->opt.Minf, &alpha); break; }   /* the reference's global alpha is set before use on every call */
    }
    flux[0] = fn[0] * F->area;
    flux[1] = (fn[1] * F->normal.c[0] + fn[2] * F->tangent1.c[0] + fn[3] * F->tangent2.c[0]) * F->area;
    flux[2] = (fn[1] * F->normal.c[1] + fn[2] * F->tangent1.c[1] + fn[3] * F->tangent2.c[1]) * F->area;
    flux[3] = (fn[1] * F->normal.c[2] + fn[2] * F->tangent1.c[2] + fn[3] * F->tangent2.c[2]) * F->area;
    flux[4] = fn[4] * F->area;
    /* SA: nu~ carried by the face mass flux, upwind face state */
    flux[5] = (o->nv == 6) ? flux[0] * ((flux[0] >= 0.) ? L->nut : R->nut) : 0.;
    if (wl) *wl = weight_left(o, L, R, fn[0]);
}

/* diffusive_face_flux (ns_diffusive_face_flux.cc:28-64), laminar */
static void diffusive_face_flux(const orc_ctx* o, const fstate* F, double flux[NVM]) {
    V3 aV = vmul(F->normal, F->area);
    V3 tx, ty, tz;
    tx.c[0] = 2. / 3. * (2. * F->gradu.c[0] - F->gradv.c[1] - F->gradw.c[2]);
    tx.c[1] = F->gradu.c[1] + F->gradv.c[0];
    tx.c[2] = F->gradu.c[2] + F->gradw.c[0];
    ty.c[0] = tx.c[1];
    ty.c[1] = 2. / 3. * (2. * F->gradv.c[1] - F->gradu.c[0] - F->gradw.c[2]);
    ty.c[2] = F->gradv.c[2] + F->gradw.c[1];
    tz.c[0] = tx.c[2];
    tz.c[1] = ty.c[2];
    tz.c[2] = 2. / 3. * (2. * F->gradw.c[2] - F->gradu.c[0] - F->gradv.c[1]);
    double mu = F->mu + F->mu_t;   /* face.mu + turb_visc (:36-39) */
    flux[1] = mu * vdot(tx, aV);
    flux[2] = mu * vdot(ty, aV);
    flux[3] = mu * vdot(tz, aV);
    flux[4] = mu * (vdot(tx, F->V) * aV.c[0] + vdot(ty, F->V) * aV.c[1] + vdot(tz, F->V) * aV.c[2]);
    double turb_cond = (o->nv == 6 || o->kw) ? o->Cp * F->mu_t / o->opt.Pr_t : 0.;   /* Cp(T_f) mu_t / Pr_t (:38) */
    double qq = (F->lambda + turb_cond) * vdot(F->gradT, aV);
    if (F->bc >= 0 && o->opt.bc[F->bc].thermal == FCFD_THERMAL_FIXED_Q) qq = o->opt.bc[F->bc].qdot * F->area;
    flux[4] += qq;
    flux[5] = (o->nv == 6) ? (F->mu + F->rho * F->nut * sa_fn(F->rho, F->nut, F->mu)) * vdot(F->gradn, aV) / sa_sigma : 0.;
}

/* ------------------------------------------------------------------ */
/* update_boundaries: decision D1                                     */
/* ------------------------------------------------------------------ */
int orc_update_boundaries(orc_ctx* o) {
    int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int f = 0; f < o->nface; ++f) {
        if (o->fbc[f] < 0 || err) continue;
        fstate F;
        face_geom_update(o, &F, f);
        int c = F.parent;
        cstate L;
        L.p = L.p_center = o->p[c];
        L.V = L.V_center = cellV(o, c);
        L.T = L.T_center = o->T[c];
        L.nut = L.nut_center = o->nut[c];
        L.rho = m_rho(o, L.p, L.T);
        L.volume = o->vol[c];
        for (int i = 0; i < NVM; ++i) L.update[i] = 0.;
        finish_state(o, &L, &F);
        cstate R = L;
        int rc = apply_bcs(o, &L, &R, &F);
        if (rc) { err |= 1; continue; }
        int g = F.neighbor;
        o->p[g] = R.p_center; o->u[g] = R.V_center.c[0]; o->v[g] = R.V_center.c[1]; o->w[g] = R.V_center.c[2]; o->T[g] = R.T_center;
        o->nut[g] = R.nut_center;
    }
    return err ? FCFD_EBC : 0;
}

/* ------------------------------------------------------------------ */
/* calc_cell_grads (src/navier_stokes/ns.cc:301-336) with
 * Variable::cell_gradient (variable.h:193-216)                        */
/* ------------------------------------------------------------------ */
static double face_value_scalar(const orc_ctx* o, const double* q, int f, int which) {
    int b = o->fbc[f];
    if (b >= 0) {
        const fcfd_bc* bc = &o->opt.bc[b];
        if (which == 0 && bc->p_fixed) return bc->p;
        if (which == 4 && bc->T_fixed) return bc->T;
        if (which >= 1 && which <= 3 && bc->V_fixed) return bc->V[which - 1];
    }
    double t = 0.;
    for (int k = o->avg_ptr[f]; k < o->avg_ptr[f + 1]; ++k) t += o->avg_w[k] * q[o->avg_idx[k]];
    return t;
}

int orc_calc_cell_grads(orc_ctx* o) {
    const double* q[6] = {o->p, o->u, o->v, o->w, o->T, o->nut};
    const int nv = o->nv;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->ncell; ++c) {
        V3 g[6];
        for (int k = 0; k < 6; ++k) g[k] = vzero();
        if (o->g_ptr[c + 1] > o->g_ptr[c]) {
            for (int e = o->g_ptr[c]; e < o->g_ptr[c + 1]; ++e) {
                int j = o->g_idx[e];
                for (int k = 0; k < nv; ++k)
                    for (int i = 0; i < 3; ++i) g[k].c[i] += o->g_w[e].c[i] * q[k][j];
            }
        } else {   /* Green-Gauss fallback */
            for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
                int f = o->cfaces[c].a[cf];
                V3 aV = vdiv(vmul(o->fn[f], o->farea[f]), o->vol[c]);
                if (o->fparent[f] != c) aV = vmul(aV, -1.);
                for (int k = 0; k < nv; ++k) {
                    double fv = face_value_scalar(o, q[k], f, k);
                    for (int i = 0; i < 3; ++i) g[k].c[i] += fv * aV.c[i];
                }
            }
        }
        o->gp[c] = g[0]; o->gu[c] = g[1]; o->gv[c] = g[2]; o->gw[c] = g[3]; o->gT[c] = g[4];
        o->gn[c] = g[5];
    }
    for (int f = 0; f < o->nface; ++f)
        if (o->fbc[f] >= 0) {
            int par = o->fparent[f], nb = o->fneigh[f];
            o->gp[nb] = o->gp[par]; o->gu[nb] = o->gu[par]; o->gv[nb] = o->gv[par]; o->gw[nb] = o->gw[par]; o->gT[nb] = o->gT[par];
            o->gn[nb] = o->gn[par];
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* limiters (src/navier_stokes/ns_limiters.cc:26-248, ns.cc:400-435)  */
/* ------------------------------------------------------------------ */
static double lim_minmod(double a1, double a2) { if (a1 * a2 < 0.) return 0.; else if (a1 > 0.) return dmin(a1, a2); else return dmax(a1, a2); }
static double isgn(double x) { return (double)((x > 0) - (x < 0)); }

int orc_calc_limiter(orc_ctx* o) {
    int lf = o->opt.limiter;
    if (lf == FCFD_LIMITER_NONE || o->opt.order == FCFD_ORDER_FIRST) return 0;
    const double* q[6] = {o->p, o->u, o->v, o->w, o->T, o->nut};
    V3* gr[6] = {o->gp, o->gu, o->gv, o->gw, o->gT, o->gn};
    const int nv = o->nv;
    double deltaRef[6] = {0, 0, 0, 0, 0, 0};
    double Lref = o->glscale, K = o->opt.limiter_threshold;
    if (lf == FCFD_LIMITER_VK) {        /* find_min_max with its index slips */
        double qmax[5], qmin[5];
        qmax[0] = qmin[0] = o->p[0]; qmax[1] = qmin[1] = o->u[0]; qmax[2] = qmin[2] = o->v[0]; qmax[3] = qmin[3] = o->w[0]; qmax[4] = qmin[4] = o->T[0];
        for (int c = 0; c < o->ncell; ++c) {
            qmax[0] = dmax(qmax[0], o->p[c]); qmax[1] = dmax(qmax[1], o->u[c]); qmax[2] = dmax(qmax[2], o->v[c]);
            qmax[3] = dmax(qmax[4], o->w[c]); qmax[4] = dmax(qmax[4], o->p[c]);
            qmin[0] = dmin(qmin[0], o->p[c]); qmin[1] = dmin(qmin[1], o->u[c]); qmin[2] = dmin(qmin[2], o->v[c]);
            qmin[3] = dmin(qmin[4], o->w[c]); qmin[4] = dmin(qmin[4], o->p[c]);
        }
        for (int i = 0; i < 5; ++i) deltaRef[i] = fabs(qmax[i] - qmin[i]);
        if (nv == 6) {   /* SA nu~: plain range (build-only variable, no reference slot) */
            double nmax = o->nut[0], nmin = o->nut[0];
            for (int c = 0; c < o->ncell; ++c) { nmax = dmax(nmax, o->nut[c]); nmin = dmin(nmin, o->nut[c]); }
            deltaRef[5] = fabs(nmax - nmin);
        }
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->ncell; ++c) {
        double phi[6] = {1., 1., 1., 1., 1., 1.};
        if (lf == FCFD_LIMITER_MINMOD) {
            const double small = 1.e-12;
            for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
                int f = o->cfaces[c].a[cf];
                V3 c2f = vsub(o->fc[f], o->cc[c]);
                int nb = (c == o->fparent[f]) ? o->fneigh[f] : o->fparent[f];
                for (int var = 0; var < nv; ++var) {
                    double dU = q[var][nb] - q[var][c];
                    double dUg = vdot(gr[var][c], c2f);
                    dU = lim_minmod(dU, dUg);
                    phi[var] = dmin(phi[var], (dU + isgn(dU) * small) / (dUg + isgn(dUg) * small));
                }
            }
        } else {
            double umax[6], umin[6];
            for (int i = 0; i < nv; ++i) umax[i] = umin[i] = q[i][c];
            for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
                int f = o->cfaces[c].a[cf];
                int nb = (c == o->fparent[f]) ? o->fneigh[f] : o->fparent[f];
                for (int i = 0; i < nv; ++i) { umax[i] = dmax(umax[i], q[i][nb]); umin[i] = dmin(umin[i], q[i][nb]); }
            }
            for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
                int f = o->cfaces[c].a[cf];
                V3 c2f = vsub(o->fc[f], o->cc[c]);
                for (int var = 0; var < nv; ++var) {
                    double dM = vdot(gr[var][c], c2f);
                    double dP = q[var][c];
                    if (lf == FCFD_LIMITER_BJ) {
                        if (dM > 0.) dP = umax[var] - dP;
                        else if (dM < 0.) dP = umin[var] - dP;
                        else { dP = 1.; dM = 1.; }
                        phi[var] = dmin(phi[var], dP / dM);
                    } else {   /* VK */
                        double eps = deltaRef[var] / Lref;
                        if (dM > 0.) dP = umax[var] - dP;
                        else if (dM < 0.) dP = umin[var] - dP;
                        double dP2 = dP * dP, dM2 = dM * dM;
                        double eps2 = pow(1.e-6 * K * eps, 2);
                        if (dM != 0.) phi[var] = dmin(phi[var], ((dP2 + eps2) + 2. * dM * dP) / (dP2 + 2. * dM2 + dM * dP + eps2));
                    }
                }
            }
        }
        for (int i = 0; i < nv; ++i) o->phi[NVM * c + i] = phi[i];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* assemble_linear_system (ns_assemble_linear_system.cc:34-267)        */
/* ------------------------------------------------------------------ */
int orc_assemble(orc_ctx* o, int with_jac) {
    int N = o->ncell;
    const int nv = o->nv, B2 = NVM * NVM;
    for (int i = 0; i < 6 * o->nbc; ++i) o->loads[i] = 0.;
    if (with_jac) {
        if (!o->JL) { o->JL = (double*)malloc(sizeof(double) * B2 * o->nface); o->JR = (double*)malloc(sizeof(double) * B2 * o->nface); o->Dg = (double*)malloc(sizeof(double) * B2 * (size_t)N); }
        memset(o->JL, 0, sizeof(double) * B2 * o->nface);
        memset(o->JR, 0, sizeof(double) * B2 * o->nface);
        memset(o->Dg, 0, sizeof(double) * B2 * (size_t)N);
    }
    const double small_number = 10. * sqrt(2.220446049250313e-16);
    const double factor = 0.001;
    /* FD step floor per variable: the reference's absolute 10 sqrt(eps) for
     * p,u,v,w,T; for SA nu~ (build-only) the same number relative to the
     * largest inlet nu~, since nu~ itself is O(1e-5) */
    double small_v[NVM];
    for (int i = 0; i < 5; ++i) small_v[i] = small_number;
    small_v[5] = small_number * sa_nut_scale(o);
    /* 1. the face loop (:75-93, get_jacobians :197-267): every face is
     *    independent, so the faces run in parallel and keep their fluxes
     *    (and FD blocks); the residual insertion follows per cell */
    int err = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
    for (int f = 0; f < o->nface; ++f) {
        if (err) continue;
        double fc[NVM], fd[NVM], fcp[NVM], fdp[NVM];
        fstate F;
        cstate L, R;
        face_geom_update(o, &F, f);
        recon_state(o, &L, &F, F.parent);
        finish_state(o, &L, &F);
        int rc = right_state_update(o, &L, &R, &F);
        if (rc) { err |= 1; continue; }
        face_state_update(o, &L, &R, &F);
        convective_face_flux(o, &L, &R, &F, fc, &o->wL[f]);
        fd[0] = 0.;
        diffusive_face_flux(o, &F, fd);
        for (int i = 0; i < NVM; ++i) { o->fC[NVM * f + i] = fc[i]; o->fD[NVM * f + i] = fd[i]; }
        if (!with_jac) continue;
        for (int var = 0; var < nv; ++var) {                           /* get_jacobians (:197-267) */
            double jl[NVM], jr[NVM] = {0, 0, 0, 0, 0, 0};
            cstate Lp = L, Rp = R;
            double eps;
            const double sn = small_v[var];
            if (L.update[var] > sn) eps = dmax(sn, factor * L.update[var]);
            else if (L.update[var] < -sn) eps = dmin(-sn, factor * L.update[var]);
            else eps = sn;
            state_perturb(o, &Lp, &F, var, eps);
            if (F.bc >= 0) { rc = right_state_update(o, &Lp, &Rp, &F); if (rc) { err |= 1; break; } }
            convective_face_flux(o, &Lp, &Rp, &F, fcp, NULL);
            face_state_adjust(o, &Lp, &Rp, &F, var);
            fdp[0] = 0.;
            diffusive_face_flux(o, &F, fdp);
            for (int j = 0; j < nv; ++j) jl[j] = (fdp[j] - fd[j] - fcp[j] + fc[j]) / eps;
            if (F.bc == -1) {
                if (R.update[var] > sn) eps = dmax(sn, factor * R.update[var]);
                else if (R.update[var] < -sn) eps = dmin(-sn, factor * R.update[var]);
                else eps = sn;
                state_perturb(o, &Rp, &F, var, eps);
                convective_face_flux(o, &L, &Rp, &F, fcp, NULL);
                face_state_adjust(o, &L, &Rp, &F, var);
                fdp[0] = 0.;
                diffusive_face_flux(o, &F, fdp);
                for (int j = 0; j < nv; ++j) jr[j] = (fdp[j] - fd[j] - fcp[j] + fc[j]) / eps;
            }
            face_state_adjust(o, &L, &R, &F, var);
            for (int j = 0; j < nv; ++j) {
                o->JL[B2 * f + NVM * var + j] = jl[j];
                o->JR[B2 * f + NVM * var + j] = jr[j];
            }
        }
    }
    if (err) return FCFD_EBC;
    /* 2. loads (:109-119), in face order */
    for (int f = 0; f < o->nface; ++f) {
        int b = o->fbc[f];
        if (b < 0 || !o->opt.bc[b].include_in_loads) continue;
        const double *fc = &o->fC[NVM * f], *fd = &o->fD[NVM * f];
        V3 t = v3(fc[1] - fd[1], fc[2] - fd[2], fc[3] - fd[3]);
        double* ld = &o->loads[6 * b];
        ld[0] += t.c[0]; ld[1] += t.c[1]; ld[2] += t.c[2];
        V3 mc = v3(o->opt.moment_center[0], o->opt.moment_center[1], o->opt.moment_center[2]);
        V3 mm = vcross(vsub(o->fc[f], mc), t);
        ld[3] += mm.c[0]; ld[4] += mm.c[1]; ld[5] += mm.c[2];
    }
    /* 3. the residual insertion (:131-140) and the diagonal blocks of the
     *    Jacobian (MatSetValues ADD, :154-177): each cell adds its faces'
     *    contributions in ascending face order, the order of the reference's
     *    face loop, so the sums round as there */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < N; ++c) {
        double r[NVM] = {0, 0, 0, 0, 0, 0}, ra[NVM] = {0, 0, 0, 0, 0, 0};
        double* D = with_jac ? &o->Dg[B2 * (size_t)c] : NULL;
        for (int k = 0; k < o->cfaces[c].n; ++k) {
            int f = o->cfaces[c].a[k];
            const double *fc = &o->fC[NVM * f], *fd = &o->fD[NVM * f];
            int parent = o->fparent[f] == c;
            for (int i = 0; i < nv; ++i) {
                double val = parent ? fd[i] - fc[i] + 0. : -1. * (fd[i] - fc[i]) + 0.;
                r[i] += val;
                ra[i] += fabs(val);
            }
            if (D) {
                for (int var = 0; var < nv; ++var)
                    for (int j = 0; j < nv; ++j) {
                        if (parent) D[NVM * j + var] += -1. * o->JL[B2 * f + NVM * var + j];
                        else D[NVM * j + var] += o->JR[B2 * f + NVM * var + j];
                    }
            }
        }
        for (int i = 0; i < NVM; ++i) { o->Rs[NVM * c + i] = r[i]; o->Rabs[NVM * c + i] = ra[i]; }
    }
    if (nv == 6) {   /* SA cell source, added after the face sums */
#pragma omp parallel for schedule(static)
        for (int c = 0; c < N; ++c) {
            double rho = m_rho(o, o->p[c], o->T[c]);
            double S = o->vol[c] * sa_source(rho, m_visc(o, o->T[c]), o->nut[c], o->gu[c], o->gv[c], o->gw[c], o->gn[c], o->wdist[c], NULL, NULL);
            o->Rs[NVM * c + 5] += S;
            o->Rabs[NVM * c + 5] += fabs(S);
            if (!with_jac) continue;
            /* its diagonal block: one-sided differences in the cell state
             * (gradients frozen) with the face perturbation policy */
            for (int var = 0; var < nv; ++var) {
                double u = o->upd[NVM * c + var], eps;
                const double sn = small_v[var];
                if (u > sn) eps = dmax(sn, factor * u);
                else if (u < -sn) eps = dmin(-sn, factor * u);
                else eps = sn;
                double q[6] = {o->p[c], o->u[c], o->v[c], o->w[c], o->T[c], o->nut[c]};
                q[var] += eps;
                double rp = m_rho(o, q[0], q[4]);
                double Sp = o->vol[c] * sa_source(rp, m_visc(o, q[4]), q[5], o->gu[c], o->gv[c], o->gw[c], o->gn[c], o->wdist[c], NULL, NULL);
                o->Dg[B2 * c + NVM * 5 + var] += -1. * ((Sp - S) / eps);
            }
        }
    }
    return 0;
}

int orc_get_residual(const orc_ctx* o, double* R, double* Rabs) {
    const int nv = o->nv;
    for (int c = 0; c < o->ncell; ++c)
        for (int i = 0; i < nv; ++i) {
            if (R) R[nv * c + i] = o->Rs[NVM * c + i];
            if (Rabs) Rabs[nv * c + i] = o->Rabs[NVM * c + i];
        }
    return 0;
}

int orc_get_fields(const orc_ctx* o, double* Wf, double* G, double* phi) {
    const int nv = o->nv;
    for (int c = 0; c < o->ntot; ++c) {
        if (Wf) {
            Wf[nv * c] = o->p[c]; Wf[nv * c + 1] = o->u[c]; Wf[nv * c + 2] = o->v[c]; Wf[nv * c + 3] = o->w[c]; Wf[nv * c + 4] = o->T[c];
            if (nv == 6) Wf[nv * c + 5] = o->nut[c];
        }
        if (G) {
            const V3* g[6] = {&o->gp[c], &o->gu[c], &o->gv[c], &o->gw[c], &o->gT[c], &o->gn[c]};
            for (int k = 0; k < nv; ++k) for (int i = 0; i < 3; ++i) G[3 * nv * c + 3 * k + i] = g[k]->c[i];
        }
    }
    if (phi) for (int c = 0; c < o->ncell; ++c) for (int i = 0; i < nv; ++i) phi[nv * c + i] = o->phi[NVM * c + i];
    return 0;
}

int orc_get_jacobian(const orc_ctx* o, double* JL, double* JR, double* Dg) {
    if (!o->JL) return FCFD_EINVAL;
    const int nv = o->nv, B2 = NVM * NVM;
    for (int f = 0; f < o->nface; ++f)
        for (int a = 0; a < nv; ++a)
            for (int b = 0; b < nv; ++b) {
                if (JL) JL[nv * nv * f + nv * a + b] = o->JL[B2 * f + NVM * a + b];
                if (JR) JR[nv * nv * f + nv * a + b] = o->JR[B2 * f + NVM * a + b];
            }
    if (Dg)
        for (int c = 0; c < o->ncell; ++c)
            for (int a = 0; a < nv; ++a)
                for (int b = 0; b < nv; ++b) Dg[nv * nv * c + nv * a + b] = o->Dg[B2 * c + NVM * a + b];
    return 0;
}

int orc_get_loads(const orc_ctx* o, double* loads) { memcpy(loads, o->loads, sizeof(double) * 6 * o->nbc); return 0; }

/* update_time_step (src/time_step.cc:143-224), no ramp / schedule */
int orc_time_step(orc_ctx* o, int dt_type, double cfl, double* dt, double* max_cfl) {
    double cur = 1.;
    if (dt_type == FCFD_DT_FIXED) {   /* time_step.cc:155-156: cfl carries the step size */
        cur = cfl;
        for (int c = 0; c < o->ncell; ++c) dt[c] = cur;
    } else if (dt_type == FCFD_DT_CFL_MAX) {
        cur = 1.e20;
        for (int c = 0; c < o->ncell; ++c) {
            double a = m_a(o, o->p[c], o->T[c]);
            cur = dmin(cur, cfl * o->lscale[c] / (vmag(cellV(o, c)) + a));
        }
        for (int c = 0; c < o->ncell; ++c) dt[c] = cur;
    } else {
        for (int c = 0; c < o->ncell; ++c) {
            double a = m_a(o, o->p[c], o->T[c]);
            dt[c] = cfl * o->lscale[c] / (vmag(cellV(o, c)) + a);
        }
    }
    double mc = -1.;
    for (int c = 0; c < o->ncell; ++c) {
        double a = m_a(o, o->p[c], o->T[c]);
        mc = dmax(mc, (vmag(cellV(o, c)) + a) * cur / o->lscale[c]);
    }
    if (dt_type == FCFD_DT_CFL_LOCAL) mc = cfl;
    *max_cfl = mc;
    return 0;
}

/* time_terms: cons2prim(c)*V/dt (ns_time_terms.cc:26-56,147-163) */
int orc_time_terms(orc_ctx* o, const double* dt, double* blocks) {
    const int nv = o->nv;
    for (int c = 0; c < o->ncell; ++c) {
        double P[6][6] = {{0}};
        double rho = m_rho(o, o->p[c], o->T[c]);
        double pressure = o->p[c] + o->opt.Pref, temperature = o->T[c] + o->opt.Tref;
        double drho_dT = -1. * rho / temperature, drho_dp = rho / pressure;
        double c_p = o->Cp;
        V3 V = cellV(o, c);
        double H = c_p * temperature + 0.5 * vdot(V, V);
        P[0][0] = drho_dp; P[0][4] = drho_dT;
        P[1][0] = drho_dp * V.c[0]; P[1][1] = rho; P[1][4] = drho_dT * V.c[0];
        P[2][0] = drho_dp * V.c[1]; P[2][2] = rho; P[2][4] = drho_dT * V.c[1];
        P[3][0] = drho_dp * V.c[2]; P[3][3] = rho; P[3][4] = drho_dT * V.c[2];
        P[4][0] = drho_dp * H - 1.; P[4][1] = rho * V.c[0]; P[4][2] = rho * V.c[1]; P[4][3] = rho * V.c[2];
        P[4][4] = drho_dT * H + rho * c_p;
        if (nv == 6) {   /* Q6 = rho nu~ */
            P[5][0] = drho_dp * o->nut[c]; P[5][4] = drho_dT * o->nut[c]; P[5][5] = rho;
        }
        for (int i = 0; i < nv; ++i) for (int j = 0; j < nv; ++j) blocks[nv * nv * c + nv * i + j] = P[i][j] * o->vol[c] / dt[c];
    }
    return 0;
}

/* probe: convective flux between two primitive states across a face with
 * normal n and tangent t1 (t2 = norm(n x t1)), as convective_face_flux;
 * out[0..4] the flux, out[5] weightL (weight_left). */
int orc_flux_probe(orc_ctx* o, int flux_type, const double* WL, const double* WR, const double* n, const double* t1,
                   double area, double* out) {
    fstate F;
    memset(&F, 0, sizeof F);
    F.normal = v3(n[0], n[1], n[2]);
    F.tangent1 = v3(t1[0], t1[1], t1[2]);
    F.tangent2 = vnorm(vcross(F.normal, F.tangent1));
    F.area = area;
    F.index = 0;
    F.bc = -1;
    cstate L, R;
    const double* src[2] = {WL, WR};
    cstate* dst[2] = {&L, &R};
    for (int k = 0; k < 2; ++k) {
        cstate* S = dst[k];
        S->p = S->p_center = src[k][0];
        S->V = S->V_center = v3(src[k][1], src[k][2], src[k][3]);
        S->T = S->T_center = src[k][4];
        S->nut = S->nut_center = 0.;
        S->rho = m_rho(o, S->p, S->T);
        finish_state(o, S, &F);
    }
    int saved = o->opt.convective_flux;
    o->opt.convective_flux = flux_type;
    double diss0 = o->fdiss ? o->fdiss[0] : 0.;
    double wd0 = o->fwd ? o->fwd[0] : 0.;
    if (o->fdiss) { o->fdiss[0] = 1.; o->fwd[0] = 1.e20; }
    double fl[NVM], wl = 0.;
    convective_face_flux(o, &L, &R, &F, fl, &wl);
    for (int i = 0; i < 5; ++i) out[i] = fl[i];
    out[5] = wl;
    if (o->fdiss) { o->fdiss[0] = diss0; o->fwd[0] = wd0; }
    o->opt.convective_flux = saved;
    return 0;
}

/* SA model pieces at one point (probe for the pin tests): out = source per
 * unit volume, P, D, mu_t, fn; gvel = du, dv, dw gradients (9) */
int orc_sa_probe(double rho, double mu, double nut, const double* gvel, const double* gn, double d, double* out) {
    V3 gu = v3(gvel[0], gvel[1], gvel[2]), gv = v3(gvel[3], gvel[4], gvel[5]), gw = v3(gvel[6], gvel[7], gvel[8]);
    double P, D;
    out[0] = sa_source(rho, mu, nut, gu, gv, gw, v3(gn[0], gn[1], gn[2]), d, &P, &D);
    out[1] = P; out[2] = D;
    out[3] = sa_mut(rho, nut, mu);
    out[4] = sa_fn(rho, nut, mu);
    return 0;
}

/* per-face fluxes of the last orc_assemble: C (convective) and D
 * (diffusive), F x nv each, reference face order */
int orc_get_face_flux(const orc_ctx* o, double* C, double* D) {
    const int nv = o->nv;
    for (int f = 0; f < o->nface; ++f)
        for (int i = 0; i < nv; ++i) {
            if (C) C[nv * f + i] = o->fC[NVM * f + i];
            if (D) D[nv * f + i] = o->fD[NVM * f + i];
        }
    return 0;
}

/* probe of apply_bcs on BC `bc` of this context: the finished left face
 * state from WL (p,u,v,w,T), its cell-centre values WLc, and the right
 * state's centre values WRc (the ghost's); the right state starts as a copy
 * of the left one, as in the face loop, and is finished after the BC.
 * out: p, u, v, w, T, rho, a, p_c, u_c, v_c, w_c, T_c of the right state.
 * Returns apply_bcs' status (FCFD_EBC where the reference exits). */
int orc_bc_probe(orc_ctx* o, int bc, const double* WL, const double* WLc, const double* WRc, const double* n,
                 const double* t1, double* out) {
    fstate F;
    memset(&F, 0, sizeof F);
    F.normal = v3(n[0], n[1], n[2]);
    F.tangent1 = v3(t1[0], t1[1], t1[2]);
    F.tangent2 = vnorm(vcross(F.normal, F.tangent1));
    F.area = 1.;
    F.bc = bc;
    cstate L, R;
    memset(&L, 0, sizeof L);
    L.p = WL[0]; L.V = v3(WL[1], WL[2], WL[3]); L.T = WL[4];
    L.p_center = WLc[0]; L.V_center = v3(WLc[1], WLc[2], WLc[3]); L.T_center = WLc[4];
    L.rho = m_rho(o, L.p, L.T);
    finish_state(o, &L, &F);
    R = L;
    R.p_center = WRc[0]; R.V_center = v3(WRc[1], WRc[2], WRc[3]); R.T_center = WRc[4];
    int rc = apply_bcs(o, &L, &R, &F);
    if (rc) return rc;
    double a_bc = R.a;
    finish_state(o, &R, &F);
    double v[12] = {R.p, R.V.c[0], R.V.c[1], R.V.c[2], R.T, R.rho, a_bc, R.p_center, R.V_center.c[0], R.V_center.c[1],
                    R.V_center.c[2], R.T_center};
    memcpy(out, v, sizeof v);
    return 0;
}

/* AUSM+-up split functions at M (probe): out = M4+, M4-, P5+(alpha), P5-(alpha) */
int orc_ausm_probe(double M, double alpha, double* out) {
    out[0] = msp4(M); out[1] = msm4(M); out[2] = psp5(M, alpha); out[3] = psm5(M, alpha);
    return 0;
}

/* Interpolate::calculate_weights on one stencil (probe for the pin tests):
 * a fresh interpolation object of dimension dim, stencil size maxs and
 * skewness tolerance tol; ids/pts (n entries, pts 3n) are sorted and
 * truncated as sort_stencil(is_internal = true); out_ids/out_w get the kept
 * ids and the weights; *kind = 4 tetra, 3 triangle, 2 line, 1 point.
 * Returns the number of weights. */
int orc_wtli_probe(int dim, int maxs, double tol, const double* point, int n, const int* ids, const double* pts,
                   int* out_ids, double* out_w, int* kind) {
    interp_t it;
    memset(&it, 0, sizeof it);
    it.dim = dim; it.maxs = maxs; it.tol = tol;
    int* id = (int*)malloc(sizeof(int) * (n + 1));
    V3* P = (V3*)malloc(sizeof(V3) * (n + 1));
    double* w = (double*)malloc(sizeof(double) * (n + 1));
    for (int i = 0; i < n; ++i) { id[i] = ids[i]; P[i] = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); }
    int nw = wtli_weights(&it, n, v3(point[0], point[1], point[2]), id, P, w, kind);
    for (int i = 0; i < nw; ++i) { out_ids[i] = id[i]; out_w[i] = w[i]; }
    free(id); free(P); free(w);
    return nw;
}

/* surface quantities of the last orc_assemble (ns_assemble_linear_system.cc:
 * 122-127): mdot.face = (C - D)[0] / A on every face; on boundary faces
 * qdot.bc = (C - D)[4] / A unless qdot is fixed on the BC (then the given
 * value) and tau.bc = -D[1..3] / A; interior faces get 0 for qdot and tau */
int orc_get_surface(const orc_ctx* o, double* mdot, double* qdot, double* tau) {
    for (int f = 0; f < o->nface; ++f) {
        const double* C = &o->fC[NVM * f];
        const double* D = &o->fD[NVM * f];
        const double A = o->farea[f];
        mdot[f] = (C[0] - D[0]) / A;
        qdot[f] = 0.; tau[3 * f] = tau[3 * f + 1] = tau[3 * f + 2] = 0.;
        int b = o->fbc[f];
        if (b >= 0) {
            qdot[f] = o->opt.bc[b].qdot_fixed ? o->opt.bc[b].qdot : (C[4] - D[4]) / A;
            for (int i = 0; i < 3; ++i) tau[3 * f + i] = -D[i + 1] / A;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* One implicit step on the host cores: the CPU side of the Newton    */
/* comparison (NavierStokes::solve, ns.cc:104-118, with the solve that */
/* the reference leaves to PETSc, ns_petsc_functions.cc:73-110, done   */
/* here as block-Jacobi right-preconditioned restarted GMRES)          */
/* ------------------------------------------------------------------ */
#include <omp.h>

/* inverse of one nv x nv block (row-major, stride NVM) by Gauss-Jordan with
 * partial pivoting; returns 1 when singular */
static int block_inverse(int nv, const double* A, double* Ainv) {
    double a[NVM][2 * NVM];
    for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j) { a[i][j] = A[NVM * i + j]; a[i][nv + j] = (i == j) ? 1. : 0.; }
    for (int c = 0; c < nv; ++c) {
        int piv = c;
        for (int r = c + 1; r < nv; ++r) if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.) return 1;
        if (piv != c) for (int j = 0; j < 2 * nv; ++j) { double t = a[c][j]; a[c][j] = a[piv][j]; a[piv][j] = t; }
        double d = a[c][c];
        for (int j = 0; j < 2 * nv; ++j) a[c][j] /= d;
        for (int r = 0; r < nv; ++r) {
            if (r == c) continue;
            double m = a[r][c];
            if (m != 0.) for (int j = 0; j < 2 * nv; ++j) a[r][j] -= m * a[c][j];
        }
    }
    for (int i = 0; i < nv; ++i) for (int j = 0; j < nv; ++j) Ainv[NVM * i + j] = a[i][nv + j];
    return 0;
}

/* y = M x with M = P V/dt - dR/dW: diagonal blocks D (time terms included)
 * and per face -JR^T (row parent, column neighbour), +JL^T (row neighbour,
 * column parent); x, y are N x nv */
static void newton_matvec(const orc_ctx* o, const double* D, const double* x, double* y) {
    const int nv = o->nv, B2 = NVM * NVM;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->ncell; ++c) {
        double acc[NVM] = {0, 0, 0, 0, 0, 0};
        const double* Dc = &D[B2 * (size_t)c];
        for (int i = 0; i < nv; ++i) for (int j = 0; j < nv; ++j) acc[i] += Dc[NVM * i + j] * x[nv * c + j];
        for (int k = 0; k < o->cfaces[c].n; ++k) {
            int f = o->cfaces[c].a[k];
            if (o->fbc[f] >= 0) continue;
            int p = o->fparent[f], n = o->fneigh[f];
            const double* Bk = (p == c) ? &o->JR[B2 * (size_t)f] : &o->JL[B2 * (size_t)f];   /* [var][eq] */
            const double sg = (p == c) ? -1. : 1.;
            const int other = (p == c) ? n : p;
            for (int eq = 0; eq < nv; ++eq) {
                double s = 0.;
                for (int var = 0; var < nv; ++var) s += Bk[NVM * var + eq] * x[nv * other + var];
                acc[eq] += sg * s;
            }
        }
        for (int i = 0; i < nv; ++i) y[nv * c + i] = acc[i];
    }
}

static void bj_apply(const orc_ctx* o, const double* Dinv, const double* x, double* y) {
    const int nv = o->nv, B2 = NVM * NVM;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->ncell; ++c)
        for (int i = 0; i < nv; ++i) {
            double s = 0.;
            for (int j = 0; j < nv; ++j) s += Dinv[B2 * (size_t)c + NVM * i + j] * x[nv * c + j];
            y[nv * c + i] = s;
        }
}

static double vdotn(size_t n, const double* a, const double* b) {
    double s = 0.;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* GMRES(m) for M x = b, right-preconditioned with block Jacobi, modified
 * Gram-Schmidt, x0 = 0; stops at ||r|| <= rtol ||b|| or max_iters */
static int gmres_bj(const orc_ctx* o, const double* D, const double* Dinv, const double* b, double* x, int m, int max_iters,
                    double rtol, double* rnorm_out) {
    const size_t n = (size_t)o->nv * o->ncell;
    double* V = (double*)malloc(sizeof(double) * n * (m + 1));
    double* z = (double*)malloc(sizeof(double) * n);
    double* w = (double*)malloc(sizeof(double) * n);
    double* H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = (double*)malloc(sizeof(double) * m), *sn = (double*)malloc(sizeof(double) * m);
    double *g = (double*)malloc(sizeof(double) * (m + 1)), *y = (double*)malloc(sizeof(double) * m);
    memset(x, 0, sizeof(double) * n);
    const double bnorm = sqrt(vdotn(n, b, b));
    double rnorm = bnorm;
    int its = 0;
    while (its < max_iters && rnorm > rtol * bnorm && bnorm > 0.) {
        /* r = b - M x */
        newton_matvec(o, D, x, w);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
        double beta = sqrt(vdotn(n, w, w));
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n; ++i) V[i] = w[i] / beta;
        for (int i = 0; i <= m; ++i) g[i] = 0.;
        g[0] = beta;
        int k = 0;
        for (; k < m && its < max_iters; ++k, ++its) {
            bj_apply(o, Dinv, &V[n * k], z);
            newton_matvec(o, D, z, w);
            for (int j = 0; j <= k; ++j) {
                double h = vdotn(n, w, &V[n * j]);
                H[(size_t)j * m + k] = h;
#pragma omp parallel for schedule(static)
                for (size_t i = 0; i < n; ++i) w[i] -= h * V[n * j + i];
            }
            double hn = sqrt(vdotn(n, w, w));
            H[(size_t)(k + 1) * m + k] = hn;
            if (hn > 0.) {
#pragma omp parallel for schedule(static)
                for (size_t i = 0; i < n; ++i) V[n * (k + 1) + i] = w[i] / hn;
            }
            for (int j = 0; j < k; ++j) {   /* previous Givens rotations */
                double a = H[(size_t)j * m + k], bb = H[(size_t)(j + 1) * m + k];
                H[(size_t)j * m + k] = cs[j] * a + sn[j] * bb;
                H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * bb;
            }
            double a = H[(size_t)k * m + k], bb = H[(size_t)(k + 1) * m + k];
            double r = sqrt(a * a + bb * bb);
            cs[k] = r > 0. ? a / r : 1.; sn[k] = r > 0. ? bb / r : 0.;
            H[(size_t)k * m + k] = r; H[(size_t)(k + 1) * m + k] = 0.;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            rnorm = fabs(g[k + 1]);
            if (rnorm <= rtol * bnorm || hn == 0.) { ++k; ++its; break; }
        }
        for (int i = k - 1; i >= 0; --i) {   /* y = H^-1 g, then x += M^-1 V y */
            double s = g[i];
            for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * y[j];
            y[i] = s / H[(size_t)i * m + i];
        }
        memset(w, 0, sizeof(double) * n);
        for (int j = 0; j < k; ++j) {
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < n; ++i) w[i] += y[j] * V[n * j + i];
        }
        bj_apply(o, Dinv, w, z);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n; ++i) x[i] += z[i];
    }
    *rnorm_out = rnorm;
    free(V); free(z); free(w); free(H); free(cs); free(sn); free(g); free(y);
    return its;
}

/* One implicit step at CFL_LOCAL `cfl` on the state held by the context:
 * residual + FD Jacobian (with the previous update setting the FD step),
 * time terms P V/dt, block-Jacobi GMRES(restart) to rtol or max_iters,
 * W += dW (kept as the next step's previous update).
 * stats: ||R||_2 before the step, Krylov iterations, then wall milliseconds
 * of residual + Jacobian, time terms + block inverses, Krylov, update. */
int orc_newton_step(orc_ctx* o, double cfl, int restart, int max_iters, double rtol, double* stats) {
    const int N = o->ncell, nv = o->nv, B2 = NVM * NVM;
    const size_t n = (size_t)nv * N;
    double t0 = omp_get_wtime();
    int rc = orc_update_boundaries(o);
    if (!rc) rc = orc_calc_cell_grads(o);
    if (!rc) rc = orc_calc_limiter(o);
    if (!rc) rc = orc_assemble(o, 1);
    if (rc) return rc;
    double* b = (double*)malloc(sizeof(double) * n);
    double rr = 0.;
    for (int c = 0; c < N; ++c) for (int i = 0; i < nv; ++i) { b[nv * c + i] = o->Rs[NVM * c + i]; rr += b[nv * c + i] * b[nv * c + i]; }
    double t1 = omp_get_wtime();
    double* dt = (double*)malloc(sizeof(double) * N);
    double mc;
    orc_time_step(o, FCFD_DT_CFL_LOCAL, cfl, dt, &mc);
    double* T = (double*)malloc(sizeof(double) * (size_t)nv * nv * N);
    orc_time_terms(o, dt, T);
    double* D = (double*)malloc(sizeof(double) * B2 * (size_t)N);
    double* Dinv = (double*)malloc(sizeof(double) * B2 * (size_t)N);
    int sing = 0;
#pragma omp parallel for schedule(static) reduction(| : sing)
    for (int c = 0; c < N; ++c) {
        double* Dc = &D[B2 * (size_t)c];
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nv; ++j) Dc[NVM * i + j] = o->Dg[B2 * (size_t)c + NVM * i + j] + T[(size_t)nv * nv * c + nv * i + j];
        sing |= block_inverse(nv, Dc, &Dinv[B2 * (size_t)c]);
    }
    double t2 = omp_get_wtime();
    double* x = (double*)malloc(sizeof(double) * n);
    double rn = 0.;
    int its = sing ? 0 : gmres_bj(o, D, Dinv, b, x, restart, max_iters, rtol, &rn);
    double t3 = omp_get_wtime();
    int bad = 0;
    for (size_t i = 0; i < n; ++i) bad |= !isfinite(x[i]);
    if (!bad && !sing) {
        double* q[6] = {o->p, o->u, o->v, o->w, o->T, o->nut};
#pragma omp parallel for schedule(static)
        for (int c = 0; c < N; ++c)
            for (int i = 0; i < nv; ++i) { q[i][c] += x[nv * c + i]; o->upd[NVM * c + i] = x[nv * c + i]; }
    }
    double t4 = omp_get_wtime();
    stats[0] = sqrt(rr); stats[1] = its;
    stats[2] = 1e3 * (t1 - t0); stats[3] = 1e3 * (t2 - t1); stats[4] = 1e3 * (t3 - t2); stats[5] = 1e3 * (t4 - t3);
    free(b); free(dt); free(T); free(D); free(Dinv); free(x);
    if (sing) { seterr(o, FCFD_EDIVERGED, "singular diagonal block"); return FCFD_EDIVERGED; }
    if (bad) { seterr(o, FCFD_EDIVERGED, "NaN/Inf in the update"); return FCFD_EDIVERGED; }
    return 0;
}

/* threads the OpenMP loops use */
int orc_num_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ */
/* dual (pseudo) time stepping (ns_time_terms.cc:58-186,               */
/* time_step.cc:226-290)                                               */
/* ------------------------------------------------------------------ */
/* cons2prim (:26-56), or with ws95 the Weiss-Smith preconditioned matrix
 * (:58-105): drho/dp replaced by 1/Ur^2 - drho/dT / (rho c_p) with the
 * reference velocity Ur = min(max(|V|, 1e-5 a), a), raised to the viscous
 * velocity mu / (rho L) when mu > 1e-8 and to 1e-5 sqrt(dpmax / rho), dpmax
 * the largest |p_c - p_n| over the cell's neighbour cells (ghosts included) */
static void pseudo_block(const orc_ctx* o, int c, int ws95, double P[6][6]) {
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) P[i][j] = 0.;
    double rho = m_rho(o, o->p[c], o->T[c]);
    double pressure = o->p[c] + o->opt.Pref, temperature = o->T[c] + o->opt.Tref;
    double drho_dT = -1. * rho / temperature, drho_dp = rho / pressure;
    double c_p = o->Cp;
    V3 V = cellV(o, c);
    if (ws95) {
        double a = m_a(o, o->p[c], o->T[c]);
        double Ur = dmax(vmag(V), 1.e-5 * a);
        Ur = dmin(Ur, a);
        double mu = m_visc(o, o->T[c]);
        if (mu > 1.e-8) Ur = dmax(Ur, mu / (rho * o->lscale[c]));
        double dpmax = 0.;
        for (int k = 0; k < o->cnbr[c].n; ++k) dpmax = dmax(dpmax, fabs(o->p[c] - o->p[o->cnbr[c].a[k]]));
        Ur = dmax(Ur, 1.e-5 * sqrt(dpmax / rho));
        drho_dp = 1. / (Ur * Ur) - drho_dT / (rho * c_p);
    }
    double H = c_p * temperature + 0.5 * vdot(V, V);
    P[0][0] = drho_dp; P[0][4] = drho_dT;
    P[1][0] = drho_dp * V.c[0]; P[1][1] = rho; P[1][4] = drho_dT * V.c[0];
    P[2][0] = drho_dp * V.c[1]; P[2][2] = rho; P[2][4] = drho_dT * V.c[1];
    P[3][0] = drho_dp * V.c[2]; P[3][3] = rho; P[3][4] = drho_dT * V.c[2];
    P[4][0] = drho_dp * H - 1.; P[4][1] = rho * V.c[0]; P[4][2] = rho * V.c[1]; P[4][3] = rho * V.c[2];
    P[4][4] = drho_dT * H + rho * c_p;
    if (o->nv == 6) { P[5][0] = drho_dp * o->nut[c]; P[5][4] = drho_dT * o->nut[c]; P[5][5] = rho; }
}

/* The implicit system of one pseudo step (time_terms with
 * pseudotime.numberofsteps > 1): blocks[N][nv][nv] = P V / dt + P_ps V / dtau
 * (P_ps: cons2prim, or the WS95 matrix) and, from the residual of the last
 * orc_assemble, rhs[N][nv] = R - (P V / dt)(W - Wn) for ps_step > 1 (R
 * alone at ps_step 1, :117-146).  dt and dtau: N each. */
int orc_pseudo_system(orc_ctx* o, const double* dt, const double* dtau, int ws95, int ps_step, const double* Wn, double* blocks,
                      double* rhs) {
    const int nv = o->nv;
    const double* q[6] = {o->p, o->u, o->v, o->w, o->T, o->nut};
    for (int c = 0; c < o->ncell; ++c) {
        double P[6][6], Pp[6][6];
        pseudo_block(o, c, 0, P);
        pseudo_block(o, c, ws95, Pp);
        for (int i = 0; i < nv; ++i) {
            double corr = 0.;
            for (int j = 0; j < nv; ++j) {
                double value = P[i][j] * o->vol[c] / dt[c];
                blocks[(size_t)nv * nv * c + nv * i + j] = value + Pp[i][j] * o->vol[c] / dtau[c];
                if (ps_step > 1) corr += value * (q[j][c] - Wn[(size_t)nv * c + j]);
            }
            rhs[(size_t)nv * c + i] = o->Rs[NVM * c + i] - corr;
        }
    }
    return 0;
}

/* ================================================================== */
/* segregated two-equation RANS (src/rans/): k-omega, k-epsilon in its */
/* omega form, BSL and SST, solved after the NS step                   */
/* ================================================================== */
typedef struct { double sigma_k, sigma_omega, beta, beta_star, kappa, alpha; } kw_coef_t;
static const kw_coef_t kw_komega = {0.85, 0.5, 0.075, 0.09, 0.41, 5. / 9.};   /* rans.cc:36-41 */
static const kw_coef_t kw_keps = {1., 0.856, 0.0828, 0.09, 0.41, 0.44};       /* rans.cc:27-32 */

static double ns_rho_cell(const orc_ctx* o, int c) { return m_rho(o, o->p[c], o->T[c]); }

/* ns.rho.face (Variable::face_calculate): the BC's rho where fixed, else the face average */
static double ns_face_rho(const orc_ctx* o, int f) {
    int b = o->fbc[f];
    if (b >= 0 && o->opt.bc[b].rho_fixed) return o->opt.bc[b].rho;
    double t = 0.;
    for (int k = o->avg_ptr[f]; k < o->avg_ptr[f + 1]; ++k) t += o->avg_w[k] * ns_rho_cell(o, o->avg_idx[k]);
    return t;
}

/* get_blending (rans_terms.cc:316-328): the F1 blending function */
static double kw_blending(double k, double omega, double rho, double y, double visc, V3 gk, V3 go) {
    double cross_diff = gk.c[0] * go.c[0] + gk.c[1] * go.c[1] + gk.c[2] * go.c[2];
    double CD = dmax(2. * rho * kw_keps.sigma_omega * cross_diff / omega, 1.e-10);
    double arg1 = 2. * sqrt(k) / (kw_keps.beta_star * omega * y);
    double arg2 = 500. * visc / (rho * y * y * omega);
    double arg3 = 4. * rho * kw_keps.sigma_omega * k / (CD * y * y);
    double arg = dmin(dmax(arg1, arg2), arg3);
    return tanh(pow(arg, 4.));
}

/* the model constants blended with F1 (rans_terms.cc:89-93, 219-223) */
static void kw_blend(double F1, kw_coef_t* c) {
    c->sigma_omega = F1 * kw_komega.sigma_omega + (1. - F1) * kw_keps.sigma_omega;
    c->sigma_k = F1 * kw_komega.sigma_k + (1. - F1) * kw_keps.sigma_k;
    c->alpha = F1 * kw_komega.alpha + (1. - F1) * kw_keps.alpha;
    c->beta = F1 * kw_komega.beta + (1. - F1) * kw_keps.beta;
    c->beta_star = F1 * kw_komega.beta_star + (1. - F1) * kw_keps.beta_star;
}

/* sqrt(2 S_ij S_ij) from the velocity gradients (rans_terms.cc:242-246) */
static double kw_strain(V3 gu, V3 gv, V3 gw) {
    double dudx = gu.c[0], dudy = gu.c[1], dudz = gu.c[2];
    double dvdx = gv.c[0], dvdy = gv.c[1], dvdz = gv.c[2];
    double dwdx = gw.c[0], dwdy = gw.c[1], dwdz = gw.c[2];
    return sqrt(2. * (dudx * dudx + dvdy * dvdy + dwdz * dwdz)
                + (dudy + dvdx) * (dudy + dvdx) + (dudz + dwdx) * (dudz + dwdx)
                + (dvdz + dwdy) * (dvdz + dwdy));
}

/* one cell of the source loop (rans_terms.cc:214-292): src[2] = the source
 * terms times the volume, jac[3] = the destruction Jacobians times the
 * volume, (k,k), (k,omega), (omega,omega); F1 = the blending (BSL, SST) or 1
 * (k-omega) or 0 (k-epsilon); sr = the strain rate */
static void kw_cell_source(int model, double F1, double rho, double k, double omega, double mu_t, double sr, double divU,
                           V3 gk, V3 go, double vol, double src[2], double jac[3]) {
    kw_coef_t m;
    kw_blend(F1, &m);
    double Prod_k = mu_t * (sr * sr - 2. / 3. * divU * divU) - 2. / 3. * rho * k * divU;
    double Dest_k = m.beta_star * rho * omega * k;
    Prod_k = dmin(Prod_k, 10. * Dest_k);
    double cross_diffusion = 0.;
    if (model != FCFD_RANS_KOMEGA) cross_diffusion = gk.c[0] * go.c[0] + gk.c[1] * go.c[1] + gk.c[2] * go.c[2];
    double s0 = Prod_k;
    s0 -= Dest_k;
    double s1 = m.alpha * rho / mu_t * Prod_k;
    s1 -= m.beta * rho * omega * omega;
    s1 += 2. * (1. - F1) * rho * kw_komega.sigma_omega * cross_diffusion / omega;
    src[0] = s0 * vol;
    src[1] = s1 * vol;
    jac[0] = m.beta_star * rho * omega * vol;
    jac[1] = m.beta_star * rho * k * vol;
    double v = 2. * m.beta * rho * omega * vol;
    v += 2. * (1. - F1) * kw_komega.sigma_omega * kw_komega.sigma_omega * cross_diffusion / (omega * omega) * vol;
    jac[2] = v;
}

/* update_boundaries of the RANS variables has no source (declared
 * rans.h:99); as for the NS ghosts, a ghost takes the BC's value where the
 * BC fixes k and omega (inlets, rans_set_bcs.cc:45) and its parent's
 * otherwise */
static void kw_update_boundaries(orc_ctx* o) {
    for (int f = 0; f < o->nface; ++f) {
        int b = o->fbc[f];
        if (b < 0) continue;
        int g = o->fneigh[f], p = o->fparent[f];
        if (o->kw_kfix[b]) { o->kk[g] = o->kw_kbc[b]; o->om[g] = o->kw_obc[b]; }
        else { o->kk[g] = o->kk[p]; o->om[g] = o->om[p]; }
    }
}

/* update_eddy_viscosity (rans_update_eddy_viscosity.cc:25-59), every cell */
static void kw_eddy_viscosity(orc_ctx* o) {
    const double a1 = 0.31;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->ntot; ++c) {
        double mu = m_visc(o, o->T[c]), rho = ns_rho_cell(o, c), d = o->wdist[c];
        double mt;
        if (o->opt.rans_model == FCFD_RANS_SST) {
            double arg1 = 2. * sqrt(o->kk[c] + 1.e-15) / (kw_keps.beta_star * o->om[c] * d);
            double arg2 = 500. * mu / (rho * o->om[c] * d * d);
            double arg3 = dmax(arg1, arg2);
            double F2 = tanh(arg3 * arg3);
            mt = a1 * rho * o->kk[c] / dmax(a1 * o->om[c], o->strain[c] * F2);
        } else {
            mt = rho * o->kk[c] / o->om[c];
        }
        o->mut[c] = dmin(o->opt.visc_ratio_limit * mu, mt);
    }
}

/* calc_cell_grads (rans.cc:172-178) with Variable::cell_gradient */
static void kw_calc_cell_grads(orc_ctx* o) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->ncell; ++c) {
        V3 a = vzero(), b = vzero();
        if (o->g_ptr[c + 1] > o->g_ptr[c]) {
            for (int e = o->g_ptr[c]; e < o->g_ptr[c + 1]; ++e) {
                int j = o->g_idx[e];
                for (int i = 0; i < 3; ++i) { a.c[i] += o->g_w[e].c[i] * o->kk[j]; b.c[i] += o->g_w[e].c[i] * o->om[j]; }
            }
        } else {   /* Green-Gauss */
            for (int cf = 0; cf < o->cfaces[c].n; ++cf) {
                int f = o->cfaces[c].a[cf];
                V3 aV = vdiv(vmul(o->fn[f], o->farea[f]), o->vol[c]);
                if (o->fparent[f] != c) aV = vmul(aV, -1.);
                double fk = kw_face_value(o, o->kk, o->kw_kfix, o->kw_kbc, f);
                double fo = kw_face_value(o, o->om, o->kw_kfix, o->kw_obc, f);
                for (int i = 0; i < 3; ++i) { a.c[i] += fk * aV.c[i]; b.c[i] += fo * aV.c[i]; }
            }
        }
        o->gk[c] = a; o->gom[c] = b;
    }
}

static void kw_refresh(orc_ctx* o) {
    kw_update_boundaries(o);
    kw_eddy_viscosity(o);
    kw_calc_cell_grads(o);
}

int orc_kw_init(orc_ctx* o, double intensity, double visc_ratio) {
    if (!o->kw) return FCFD_EINVAL;
    /* apply_initial_conditions (rans.cc:120-170), one region over the whole domain */
    for (int c = 0; c < o->ncell; ++c) {
        double k = intensity * vmag(cellV(o, c));
        k *= 1.5 * k;
        k = dmax(o->opt.k_low, k);
        o->kk[c] = k;
        o->om[c] = ns_rho_cell(o, c) * k / (visc_ratio * m_visc(o, o->T[c]));
        o->gk[c] = vzero(); o->gom[c] = vzero();
    }
    for (int c = 0; c < o->ntot; ++c) o->strain[c] = 0.;
    /* set_bcs (rans_set_bcs.cc:39-73): inlet values from the NS face values
     * of the BC's first face */
    for (int b = 0; b < o->nbc; ++b) {
        const fcfd_bc* bc = &o->opt.bc[b];
        if (bc->type != FCFD_BC_INLET) continue;
        int fid = 0;
        for (int f = 0; f < o->nface; ++f) if (o->fbc[f] == b) { fid = f; break; }
        double rho = ns_face_rho(o, fid);
        double T = face_value_scalar(o, o->T, fid, 4);
        V3 V = v3(face_value_scalar(o, o->u, fid, 1), face_value_scalar(o, o->v, fid, 2), face_value_scalar(o, o->w, fid, 3));
        double kb = bc->turb_intensity * vmag(V);
        kb *= 1.5 * kb;
        kb = dmax(o->opt.k_low, kb);
        o->kw_kbc[b] = kb;
        o->kw_obc[b] = rho * kb / (bc->eddy_visc_ratio * m_visc(o, T));
        o->kw_mbc[b] = bc->eddy_visc_ratio * m_visc(o, T);
    }
    kw_refresh(o);
    return 0;
}

/* k, omega of the interior cells (N x 2), then ghosts, mu_t and gradients */
int orc_kw_set(orc_ctx* o, const double* K) {
    if (!o->kw) return FCFD_EINVAL;
    for (int c = 0; c < o->ncell; ++c) { o->kk[c] = K[2 * c]; o->om[c] = K[2 * c + 1]; }
    kw_refresh(o);
    return 0;
}

/* the inlet values k, omega, mu_t of BC b (kbc[3]) */
int orc_kw_bc_values(const orc_ctx* o, int b, double* out) {
    if (b < 0 || b >= o->nbc) return FCFD_EINVAL;
    out[0] = o->kw_kbc[b]; out[1] = o->kw_obc[b]; out[2] = o->kw_mbc[b];
    return 0;
}

/* RANS::terms + time_terms (rans_terms.cc:39-292, rans_time_terms.cc:25-86
 * without pseudo time) after orc_assemble at the same NS state: rhs[N][2],
 * diagonal blocks diag[N][2][2] (row-major; every face's, then the sources',
 * then rho V / dt), per face jf[F][6] = the face's contributions to its
 * parent's rhs (diffusive - convective) for k and omega, jacL_k,
 * jacL_omega, jacR_k, jacR_omega (jacR 0 on boundary faces; the operator's
 * row P, column N entry is jacR, row N, column P is -jacL).  Sets the
 * strain rate of every cell. */
int orc_kw_terms(orc_ctx* o, const double* dt, double* rhs, double* diag, double* jf) {
    if (!o->kw) return FCFD_EINVAL;
    const int F = o->nface, N = o->ncell, model = o->opt.rans_model;
    const int blended = model == FCFD_RANS_BSL || model == FCFD_RANS_SST;
    double* fv = (double*)malloc(sizeof(double) * 6 * (size_t)F);
#pragma omp parallel for schedule(static)
    for (int f = 0; f < F; ++f) {
        double blending = (model == FCFD_RANS_KOMEGA) ? 1. : 0.;
        int parent = o->fparent[f], neighbor = o->fneigh[f], b = o->fbc[f];
        double lam_visc = m_visc(o, face_value_scalar(o, o->T, f, 4));
        double turb_visc = kw_face_value(o, o->mut, o->kw_mfix, o->kw_mbc, f);
        double area = o->farea[f];
        double mdot = (o->fC[NVM * f] - o->fD[NVM * f]) / area;   /* mdot.face (ns_assemble_linear_system.cc:122) */
        double weightL = o->wL[f], weightR = 1. - weightL;
        /* get_kOmega (rans_terms.cc:271-314) */
        double leftK = o->kk[parent], leftOmega = o->om[parent], rightK = o->kk[neighbor], rightOmega = o->om[neighbor];
        int extrapolated = 0;
        V3 faceGradK, faceGradOmega;
        if (b >= 0) {
            int ty = o->opt.bc[b].type;
            if (ty == FCFD_BC_SYMMETRY || ty == FCFD_BC_OUTLET) extrapolated = 1;
            else if (ty == FCFD_BC_WALL && o->opt.bc[b].kind == FCFD_KIND_SLIP) extrapolated = 1;
            faceGradK = o->gk[parent]; faceGradOmega = o->gom[parent];
        } else {
            faceGradK = vzero(); faceGradOmega = vzero();
            for (int k = o->avg_ptr[f]; k < o->avg_ptr[f + 1]; ++k) {
                faceGradK = vadd(faceGradK, vsc(o->avg_w[k], o->gk[o->avg_idx[k]]));
                faceGradOmega = vadd(faceGradOmega, vsc(o->avg_w[k], o->gom[o->avg_idx[k]]));
            }
        }
        double faceK = kw_face_value(o, o->kk, o->kw_kfix, o->kw_kbc, f);
        double faceOmega = kw_face_value(o, o->om, o->kw_kfix, o->kw_obc, f);
        double faceRho = ns_face_rho(o, f);
        V3 left2right = vsub(o->cc[neighbor], o->cc[parent]);
        V3 l2rn = vnorm(left2right);
        double l2rmag = vmag(left2right);
        faceGradK = vsub(faceGradK, vsc(vdot(faceGradK, l2rn), l2rn));
        faceGradK = vadd(faceGradK, vsc((rightK - leftK) / l2rmag, l2rn));
        faceGradOmega = vsub(faceGradOmega, vsc(vdot(faceGradOmega, l2rn), l2rn));
        faceGradOmega = vadd(faceGradOmega, vsc((rightOmega - leftOmega) / l2rmag, l2rn));
        double conv0 = mdot * (weightL * leftK + weightR * rightK) * area;
        double conv1 = mdot * (weightL * leftOmega + weightR * rightOmega) * area;
        if (blended) blending = kw_blending(faceK, faceOmega, faceRho, o->wdist[parent], lam_visc, faceGradK, faceGradOmega);
        kw_coef_t m;
        kw_blend(blending, &m);
        double dn = vdot(faceGradK, o->fn[f]), dno = vdot(faceGradOmega, o->fn[f]);
        double diff0 = (lam_visc + turb_visc * m.sigma_k) * dn * area;
        double diff1 = (lam_visc + turb_visc * m.sigma_omega) * dno * area;
        double AoverH = area * vdot(o->fn[f], vnorm(left2right)) / l2rmag;
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
        double* q = &fv[6 * (size_t)f];
        q[0] = diff0 - conv0; q[1] = diff1 - conv1; q[2] = jl0; q[3] = jl1; q[4] = jr0; q[5] = jr1;
        if (jf) for (int i = 0; i < 6; ++i) jf[6 * (size_t)f + i] = q[i];
        if (f == F - 1) o->kw_lam_last = lam_visc;
    }
    /* the strain rate of every cell (rans_terms.cc:242-249) */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->ntot; ++c) o->strain[c] = kw_strain(o->gu[c], o->gv[c], o->gw[c]);
    /* per cell: faces in ascending order (the order of the face loop's
     * ADD_VALUES), then the sources, then the time term */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < N; ++c) {
        double r0 = 0., r1 = 0., d00 = 0., d11 = 0.;
        for (int k = 0; k < o->cfaces[c].n; ++k) {
            int f = o->cfaces[c].a[k];
            const double* q = &fv[6 * (size_t)f];
            if (o->fparent[f] == c) { r0 += q[0]; r1 += q[1]; d00 += q[2]; d11 += q[3]; }
            else { r0 += -1. * q[0]; r1 += -1. * q[1]; d00 += -q[4]; d11 += -q[5]; }
        }
        double rho = ns_rho_cell(o, c);
        double F1 = (model == FCFD_RANS_KOMEGA) ? 1. : 0.;
        if (blended) F1 = kw_blending(o->kk[c], o->om[c], rho, o->wdist[c], o->kw_lam_last, o->gk[c], o->gom[c]);
        double divU = o->gu[c].c[0] + o->gv[c].c[1] + o->gw[c].c[2];
        double src[2], jac[3];
        kw_cell_source(model, F1, rho, o->kk[c], o->om[c], o->mut[c], o->strain[c], divU, o->gk[c], o->gom[c], o->vol[c], src, jac);
        r0 += src[0]; r1 += src[1];
        d00 += jac[0]; d11 += jac[2];
        double tt = rho * o->vol[c] / dt[c];   /* rans_time_terms.cc:65-66 */
        d00 += tt; d11 += tt;
        rhs[2 * c] = r0; rhs[2 * c + 1] = r1;
        diag[4 * c] = d00; diag[4 * c + 1] = jac[1]; diag[4 * c + 2] = 0.; diag[4 * c + 3] = d11;
    }
    free(fv);
    return 0;
}

/* update_variables (rans.cc:180-252) with the update dU[N][2] and the NS
 * time step dt[N]: the limits, k += dk, omega += domega, then the ghosts,
 * mu_t (with the strain rate of the last orc_kw_terms) and the gradients.
 * sums[2] = sum dk^2/dt^2, sum domega^2/dt^2.  FCFD_EDIVERGED on a
 * NaN/Inf update (nothing changed). */
int orc_kw_update(orc_ctx* o, const double* dU, const double* dt, double* sums) {
    if (!o->kw) return FCFD_EINVAL;
    for (int c = 0; c < o->ncell; ++c)
        for (int i = 0; i < 2; ++i)
            if (isnan(dU[2 * c + i]) || isinf(dU[2 * c + i])) return FCFD_EDIVERGED;
    double s0 = 0., s1 = 0.;
    for (int c = 0; c < o->ncell; ++c) {
        double u0 = dU[2 * c], u1 = dU[2 * c + 1];
        u0 = dmax(-1. * (o->kk[c] - o->opt.k_low), u0);
        u0 = dmin((o->opt.k_high - o->kk[c]), u0);
        u1 = dmax(-1. * (o->om[c] - o->opt.omega_low), u1);
        o->kk[c] += u0;
        o->om[c] += u1;
        double dt2 = dt[c] * dt[c];
        s0 += u0 * u0 / dt2;
        s1 += u1 * u1 / dt2;
    }
    if (sums) { sums[0] = s0; sums[1] = s1; }
    kw_refresh(o);
    return 0;
}

/* K[ntot][2], mu_t[ntot], G[ntot][6] (grad k, grad omega), strain[ntot],
 * mutf[F] (mu_t.face); any may be NULL */
int orc_kw_get(const orc_ctx* o, double* K, double* mut, double* G, double* strain, double* mutf) {
    for (int c = 0; c < o->ntot; ++c) {
        if (K) { K[2 * c] = o->kk[c]; K[2 * c + 1] = o->om[c]; }
        if (mut) mut[c] = o->mut[c];
        if (strain) strain[c] = o->strain[c];
        if (G) for (int i = 0; i < 3; ++i) { G[6 * c + i] = o->gk[c].c[i]; G[6 * c + 3 + i] = o->gom[c].c[i]; }
    }
    if (mutf) for (int f = 0; f < o->nface; ++f) mutf[f] = kw_face_value(o, o->mut, o->kw_mfix, o->kw_mbc, f);
    return 0;
}

/* weightL of the faces of the last orc_assemble (F) */
int orc_get_weight_left(const orc_ctx* o, double* wl) { memcpy(wl, o->wL, sizeof(double) * o->nface); return 0; }

/* one cell's source terms for the pin tests: gvel = du, dv, dw gradients
 * (9), gk, go (3 each); lam_visc enters the blending; out = source k,
 * source omega, Jacobians (k,k), (k,omega), (omega,omega), F1, strain rate,
 * SST/k-omega mu_t of the point (the strain rate's) */
int orc_kw_source_probe(int model, double rho, double k, double omega, double mu_t, double lam_visc, const double* gvel,
                        const double* gk, const double* go, double d, double vol, double* out) {
    V3 gu = v3(gvel[0], gvel[1], gvel[2]), gv = v3(gvel[3], gvel[4], gvel[5]), gw = v3(gvel[6], gvel[7], gvel[8]);
    V3 K = v3(gk[0], gk[1], gk[2]), O = v3(go[0], go[1], go[2]);
    double F1 = (model == FCFD_RANS_KOMEGA) ? 1. : 0.;
    if (model == FCFD_RANS_BSL || model == FCFD_RANS_SST) F1 = kw_blending(k, omega, rho, d, lam_visc, K, O);
    double sr = kw_strain(gu, gv, gw);
    double divU = gu.c[0] + gv.c[1] + gw.c[2];
    double src[2], jac[3];
    kw_cell_source(model, F1, rho, k, omega, mu_t, sr, divU, K, O, vol, src, jac);
    out[0] = src[0]; out[1] = src[1]; out[2] = jac[0]; out[3] = jac[1]; out[4] = jac[2]; out[5] = F1; out[6] = sr;
    return 0;
}
