// Device-side mesh layout (internal numbering) and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "fcfd_internal.h"

namespace fcfd {

// All arrays are device pointers; cells and faces are in internal (space-filling-curve or RCM)
// order.  Cell arrays are padded to `ld` entries; interior cells [0,N), ghosts
// [N, N+B).  Per-cell variable arrays are [nv][ld] (p, u, v, w, T, and nu~
// with the SA model), gradients 3 nv and limiters nv components per cell in
// row pairs (pidx below), face fluxes [nv][F].  (Structure of arrays: consecutive faces have consecutive parent
// cells, so a wavefront's parent-side gathers coalesce; per-cell records and
// face-major flux rows were measured slower.)  A half-face entry hf_face[h] is the internal face id f when the
// owning cell is the face's parent and ~f when it is the neighbour; the
// half-faces of a cell are in ascending reference face order, so summing them
// in order reproduces the reference's VecSetValues(ADD) sequence.
struct DevGrid {
    int N, B, F, ld, nhf;
    int nv;                             // variables per cell (5, or 6 with SA)
    const double* wd;                   // [ld] closest wall distance of each cell (SA)
    int ref_first, ref_last;            // internal ids of reference cells 0 and N-1
    const double *ccx, *ccy, *ccz, *vol, *lscale;
    const int* gslice;                  // N/64+1 entry offsets of 64-wide gradient slices
    const double2 *gwa, *gwb;           // per entry (w_x, w_y) and (w_z, neighbour index as int64 bits)
    const int *hf_ptr, *hf_face;        // N+1, nhf
    const int* hf_nb;                   // nhf: the cell across each half-face (neighbour or parent; a ghost for a boundary face)
    const int *fpar, *fnei, *fbc;
    const double *fnx, *fny, *fnz, *t1x, *t1y, *t1z, *farea, *fcx, *fcy, *fcz, *fwp, *fwn, *fwd, *fdiss;
    const int* gh_face;                 // B: boundary face of ghost N+i
    const BcDev* bc;
    // the same face data packed for the face kernels (one 16-byte load per
    // pair): fcell[f] = (parent, neighbour, bc, 0); fgeo [6][F] row pairs
    // (n.x, n.y) (n.z, t1.x) (t1.y, t1.z) (area, c.x) (c.y, c.z) (w_parent, w_neighbour)
    const int4* fcell;
    const double2* fgeo;
    const double2* cc2;                 // [2][ld] cell centroid and volume packed: (x, y), (z, V)
    // Green-Gauss gradients (NULL when every cell has a gradient map):
    // ggc[i] = 1 for a cell without a map; ggw[f] = the face-average weights
    // (parent, neighbour or ghost) of internal face f
    const int* ggc;
    const double2* ggw;
    // eddy viscosity of the segregated two-equation model per cell ([ld],
    // interior and ghosts; NULL without it), read by the VISC = kViscRans face
    // kernels
    const double* kw_mut;
    // face-average stencils wider than the face's two cells (WTLI or IDW with
    // a stencil above 2, viscous cases; wslot NULL when every interior face
    // averages over its parent and neighbour).  wslot[f] = the slot of
    // internal face f (-1: two-cell); slot s averages entries wptr[s] ..
    // wptr[s+1] (cells as internal ids, weights) in the reference's key order.
    // Its face gradients of u, v, w, T (and nu~), 3 (nv-1) rows [.][nwide],
    // are formed once per gradient pass (launch_wide_grad): wgrad from G,
    // wgrad_d / wgrad_f the FP64 / FP32 tangent lanes' from dG.
    const int* wslot;
    int nwide;
    const int* wptr;
    const int* widx;
    const double* ww;
    double* wgrad;
    double* wgrad_d;
    float* wgrad_f;
    // the same stencils transposed, for the adjoint: interior cell c appears
    // in the wide faces wtface[wtptr[c] .. wtptr[c+1]) (internal ids, ascending)
    // with the weights wtw
    const int* wtptr;
    const int* wtface;
    const double* wtw;
};

// Gradients ([3 nv] components per cell), limiters ([nv]) and the FD
// Jacobian blocks (nv^2 entries per face or cell, FP64 and FP32) are stored in
// row pairs: component k of cell i at ((k / 2) ld + i) 2 + k % 2, the row
// count rounded up to even (the pad row holds 0).  A lane then loads two
// components of its cell with one 16-byte load while the lanes of a
// wavefront still read adjacent cells, which halves the load instructions of
// the gather kernels (bound by their memory pipeline, not by bytes).
__host__ __device__ constexpr int pair_groups(int rows) { return (rows + 1) / 2; }
__host__ __device__ __forceinline__ size_t pidx(size_t ld, int k, size_t i) { return ((size_t)(k >> 1) * ld + i) * 2 + (k & 1); }

// entries between consecutive contiguous nv x nv blocks: nv^2 rounded up to
// even, so that every block starts on a 16-byte boundary
__host__ __device__ constexpr int block_stride(int nv) { return 2 * pair_groups(nv * nv); }

template <typename T> struct PairOf;
template <> struct PairOf<double> { using type = double2; };
template <> struct PairOf<float> { using type = float2; };

template <int K, typename T>
__device__ __forceinline__ void store_pairs(T* __restrict__ A, size_t ld, size_t i, const T (&v)[K]) {
    using P2 = typename PairOf<T>::type;
    P2* A2 = reinterpret_cast<P2*>(A);
#pragma unroll
    for (int j = 0; j < pair_groups(K); ++j) {
        P2 t;
        t.x = v[2 * j];
        t.y = 2 * j + 1 < K ? v[2 * j + 1] : T(0);
        A2[(size_t)j * ld + i] = t;
    }
}
template <int K, typename T>
__device__ __forceinline__ void load_pairs(const T* __restrict__ A, size_t ld, size_t i, T (&v)[K]) {
    using P2 = typename PairOf<T>::type;
    const P2* A2 = reinterpret_cast<const P2*>(A);
#pragma unroll
    for (int j = 0; j < pair_groups(K); ++j) {
        const P2 t = A2[(size_t)j * ld + i];
        v[2 * j] = t.x;
        if (2 * j + 1 < K) v[2 * j + 1] = t.y;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2: workgroup b runs on XCD b % 8.  xcd_block() renumbers them so that XCD x
// takes the contiguous range [x q + min(x, r), ...) of logical blocks (q, r =
// gridDim / 8, gridDim % 8): faces and cells that are close in the internal
// numbering, and so share gathered cell data, then share an L2.  A bijection
// of [0, gridDim) for every grid size.
// threads per workgroup of the per-cell kernels: 256, or 64 below 131,072
// cells (fewer than 512 workgroups of 256), so that a small mesh or a coarse
// multigrid level spreads over more CUs (a 32K-cell Newton race ran 16%
// faster); large meshes keep 256
inline unsigned cell_threads(size_t n) { return n < 131072 ? 64u : 256u; }
inline dim3 cell_grid(size_t n) {
    const unsigned t = cell_threads(n);
    return dim3((unsigned)((n + t - 1) / t));
}

// faces of a cell: at most 6 for the supported cell types (tetrahedron,
// pyramid, prism, hexahedron); the set-up rejects meshes with more
constexpr int kMaxCellFaces = 6;

constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int n = gridDim.x, b = blockIdx.x;
    const int q = n / kXcds, r = n % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

void launch_ghosts(const DevGrid& g, const Params& P, double* W, int* status, hipStream_t s);
// gradients and limiters; with the SA model and src != nullptr also V_c S_c
// (the nu~ source of each interior cell, from the fresh gradients) into src[ld]
void launch_grad_lim(const DevGrid& g, const Params& P, const double* W, double* G, double* PHI, double* dref, hipStream_t s,
                     double* src = nullptr);
// face fluxes D - C per face into FX ([nv][F]) ...
void launch_flux(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, double* FX,
                 double* R, int* status, hipStream_t s);
// surface quantities of the face loop into out [5][F]: mdot (every face),
// qdot, tau_x, tau_y, tau_z (boundary faces; 0 on interior faces)
void launch_surface(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI, double* out, int* status,
                    hipStream_t s);
// ... and their ordered per-cell sums into R (no source: tangents)
void launch_cell_sum(const DevGrid& g, const double* FX, double* R, hipStream_t s);
void launch_cell_sum(const DevGrid& g, const float* FX, float* R, hipStream_t s);   // FP32 tangent sums
// the residual's per-cell sums, plus V S (src, from launch_grad_lim) in the
// nu~ row with the SA model; the FP32 form adds the FP32 tangent's source
void launch_residual_sum(const DevGrid& g, const double* src, const double* FX, double* R, hipStream_t s);
void launch_residual_sum(const DevGrid& g, const float* src, const float* FX, float* R, hipStream_t s);
// the convective flux of a batch of faces from given left/right primitive
// states (SoA [5][n]), normals and first tangents ([3][n]), areas; diss and
// wdist (Steger-Warming only) may be null: 1 and 1e20
void launch_flux_leaf(const Params& P, int flux, int n, const double* WL, const double* WR, const double* nrm, const double* tan1,
                      const double* area, const double* diss, const double* wdist, double* out, double* wl, hipStream_t s);
// the wide face-average stencils' face gradients (g.wslot set): out[k][s] =
// sum over the entries of slot s, from 0 in key order, of w * G row 3 + k
// (u, v, w, T, nu~ gradient components) of the entry's cell
void launch_wide_grad(const DevGrid& g, const double* G, double* out, hipStream_t s);
void launch_wide_grad(const DevGrid& g, const float* G, float* out, hipStream_t s);
// VK reference range only (find_min_max), when the limiter threshold needs it
void launch_vk_range(const DevGrid& g, const Params& P, const double* W, double* dref, hipStream_t s);
constexpr int kDrefSize = 8 + 10 * 1024;   // dref[nv] + VK min/max partials
void launch_dt_local(const DevGrid& g, const Params& P, const double* W, double cfl, double* dt, hipStream_t s);
// Green-Gauss gradients of the cells without a gradient map into G (row
// pairs); tangent = 1 for a tangent state (BC-fixed face values give 0).
// No-op when every cell has a map.
void launch_green_gauss(const DevGrid& g, const double* W, double* G, int tangent, hipStream_t s, const BcSeed* seed = nullptr);
// dt of every cell for FCFD_DT_FIXED (value = the step size), FCFD_DT_CFL_MAX
// (the minimum over the cells of value L / (|V| + a), kept in red[0]) or
// FCFD_DT_CFL_LOCAL (value = the CFL number); red holds 2 + 256 doubles
void launch_time_step(const DevGrid& g, const Params& P, const double* W, int dt_type, double value, double* dt, double* red,
                      hipStream_t s);
// *dst = *src, or val when src is NULL (one thread, stream-ordered)
void launch_store_scalar(const double* src, double val, double* dst, hipStream_t s);
// red[1] = the largest (|V| + a) step / L over the cells, step = *step or step_val
void launch_max_cfl(const DevGrid& g, const Params& P, const double* W, const double* step, double step_val, double* red,
                    hipStream_t s);
// brute-force wall distances (cells: T centroids; faces: with the SW dissipation factor)
void launch_wall_distance(int T, const double* ccx, const double* ccy, const double* ccz, int nw, const double* wx, const double* wy,
                          const double* wz, double* wd, hipStream_t s);
void launch_face_wall(int F, const double* fcx, const double* fcy, const double* fcz, const double* fnx, const double* fny,
                      const double* fnz, const int* is_wall, int nw, const double* wx, const double* wy, const double* wz,
                      const double* wnx, const double* wny, const double* wnz, double* fwd, double* fdiss, hipStream_t s);

}  // namespace fcfd
