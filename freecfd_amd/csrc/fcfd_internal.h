// Internal declarations shared by the host set-up, the device kernels and the
// C-ABI entry points of libfcfd.  Not part of the public interface.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fcfd_gpu.h"

namespace fcfd {

// Mesh + metrics in reference numbering, built on the host by fcfd_setup.cpp.
struct HostGrid {
    int dim = 3;
    int nnode = 0, N = 0, B = 0, F = 0;
    std::vector<double> node;                 // 3*nnode (local node order)
    std::vector<int> cell_nptr, cell_nodes;   // N+1, nodes per interior cell (local ids)
    std::vector<int> fparent, fneigh, fbc;    // F
    std::vector<int> fnptr, fnodes;           // F+1, face node lists
    std::vector<double> fc, fn, ft1, farea;   // 3F, 3F, 3F, F
    std::vector<int> fsym, face2bc;           // F
    std::vector<double> fwd, fdiss;           // F: closest wall distance, SW dissipation factor
    std::vector<double> cc, vol, lscale, wdist;   // 3(N+B), N+B, N+B, N+B
    std::vector<int> cbc;                     // N+B (ghost's BC, -1 interior)
    std::vector<int> ghost_face;              // B: the boundary face of each ghost
    std::vector<int> cfptr, cfaces;           // N+1, faces of interior cells, ascending
    std::vector<int> nbptr, nb;               // N+B+1, neighborCells
    std::vector<int> gptr, gidx;              // N+1, gradient map keys (ascending)
    std::vector<double> gw;                   // 3*entries
    std::vector<int> aptr, aidx;              // F+1, face-average keys (ascending)
    std::vector<double> aw;
    std::vector<int> naptr, naidx;            // nnode+1, node-average keys (ascending cells, ghosts included)
    std::vector<double> naw;
    std::vector<double> bc_area;              // per BC total area
    double total_volume = 0.0, glscale = 0.0;
    // internal numbering
    std::vector<int> perm;                    // N+B: internal id of reference cell
    std::vector<int> iperm;                   // N+B: reference id of internal cell
    std::vector<int> fperm;                   // F: internal face id of reference face
    std::vector<int> ifperm;                  // F: reference face of internal face
};

int build_host_grid(const fcfd_mesh* m, const fcfd_options* o, HostGrid& g, std::string& err,
                    bool wall_distance_on_host);
// internal cell numbering: reverse Cuthill-McKee or space-filling curve
constexpr int kOrderRcm = 0, kOrderSfc = 1;
void renumber(HostGrid& g, int cell_order);

// A direction in the values of one BC (index; -1: none), for the exact
// derivatives of the residual and the loads with respect to BC values.
struct BcSeed {
    int index;
    double p, T, rho, V[3], p_total, T_total, mdot, qdot, nut;
};

// Material/numerics constants passed by value to kernels.
struct Params {
    double gamma, R, Cp, Pref, Tref;
    int visc_model; double mu, sut_mu_ref, sut_T_ref, sut_S;
    int cond_model; double lambda, Pr;
    int order, limiter, flux;
    double limiter_threshold, Minf, wdiss, bl_height;
    double glscale;
    int viscous;     // 0: mu and lambda identically zero -> diffusive flux is exactly 0
    int nv;          // variables per cell: 5, or 6 with the SA model (nu~ last)
    double Pr_t;     // turbulent Prandtl number (SA)
    double nut_eps;  // FD step floor for nu~ (10 sqrt(eps) times the largest inlet nu~)
    // segregated two-equation RANS (turbulence = FCFD_TURB_KW): the NS face
    // loop takes mu_t from the model; model constants and limits
    int kw, rans_model;
    double k_low, k_high, omega_low, visc_ratio_limit;
    BcSeed seed;     // the parameter tangent's direction (index -1 elsewhere)
};

// the face kernels' VISC template argument: 0 inviscid, 1 viscous (laminar or
// SA), 2 viscous with the eddy viscosity of the segregated two-equation model
constexpr int kViscRans = 2;

// BC table entry on the device (uniform BC values).
struct BcDev {
    int type, kind, specified, thermal;
    double p, T, rho, V[3], p_total, T_total, mdot, qdot, total_area;
    double nut;      // SA nu~ prescribed at an inlet
    int p_fixed, T_fixed, V_fixed;   // Variable::fixedonBC of p, T, V (face values in Green-Gauss gradients)
    int qdot_fixed;                  // qdot.fixedonBC: the face loop keeps the given qdot (ns_assemble_linear_system.cc:125)
    int rho_fixed;                   // rho.fixedonBC (ns.rho.face of the two-equation model's face loop)
    // two-equation RANS face values (rans_set_bcs.cc:39-79): k and omega fixed
    // at inlets, mu_t fixed at inlets and no-slip walls (0 there)
    int kw_kfix, kw_mfix;
    double kw_k, kw_om, kw_mut;
    double kw_turb_intensity, kw_eddy_visc_ratio;   // the inlet's turbulenceintensity, eddyviscosityratio
};

}  // namespace fcfd
