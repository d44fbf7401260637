// Aggregation multigrid preconditioner on the FD Jacobian blocks of the
// implicit operator (fcfd_mg.hip).  No counterpart in the reference, whose
// linear solve is PETSc's (ns_petsc_functions.cc:25-110); this is the
// build's coarse-space complement to block Jacobi / block SGS.
//
// Hierarchy: level 0 is the mesh (interior cells, face adjacency, the FD
// blocks D, JL, JR); level l+1 aggregates level l by plain aggregation (a
// seed cell with all its still-free neighbours, then every leftover cell
// joins an aggregate next to it).  Coarse operators are Galerkin products
// with piecewise-constant prolongation, A_{l+1}[I][J] = sum over i in I,
// j in J of A_l[i][j]: each coarse block is a fixed, host-ordered list of
// finer blocks, summed by one thread (deterministic), re-assembled on every
// implicit step from the new FD blocks.  The V-cycle runs in FP64: multicolour
// block Gauss-Seidel pre-smoothing (colours in forward order) and
// post-smoothing (backward order), residual, restriction (sum over the
// aggregate), coarse correction, prolongation (injection); symmetric sweeps
// alone on the coarsest level.  A fixed linear operator, so GMRES may use it as a
// right preconditioner.
#pragma once
#include <vector>

#include "fcfd_device.h"
#include "fcfd_krylov.h"

namespace fcfd {

struct MgLevel {
    int n = 0, ld = 0, nnz = 0;          // cells, padded vector stride, off-diagonal blocks
    // link to the finer level: aggregate of each finer cell, members of each aggregate
    const int* agg = nullptr;            // [n_finer]
    const int *mptr = nullptr, *mem = nullptr;   // [n+1], [n_finer]
    // block rows without the diagonal: columns of row I are col[rptr[I] .. rptr[I+1])
    const int *rptr = nullptr, *col = nullptr;
    std::vector<int> rptr_h;
    // the transposed pattern: row I of A^T holds, for J in tcol[tptr[I] ..
    // tptr[I+1]) ascending, the block A(J, I) stored at tidx (transposed on use)
    const int *tptr = nullptr, *tcol = nullptr, *tidx = nullptr;
    // Galerkin assembly lists: diagonal block I sums dsrc[dptr[I] .. dptr[I+1]),
    // off-diagonal block k sums osrc[optr[k] .. optr[k+1]) (codes: see fcfd_mg.hip)
    const int *dptr = nullptr, *dsrc = nullptr, *optr = nullptr, *osrc = nullptr;
    double *D = nullptr, *A = nullptr, *Dinv = nullptr;   // n, nnz, n contiguous blocks of block_stride(nv) entries
    double *b = nullptr, *x = nullptr, *r = nullptr;      // [nv][ld]
    Colors colors;
    const int* cptr_dev = nullptr;       // colours.ptr on the device (small levels' one-launch sweeps)
    bool one_launch = false;             // sweeps of every colour in one workgroup launch (small levels)
};

struct Multigrid {
    bool ready = false;
    int nv = 0;
    std::vector<MgLevel> lv;             // lv[0] describes the mesh level (only n, ld, b/x/r used)
    double *fine_r = nullptr;            // [nv][ld] fine residual
    double *fine_in = nullptr, *fine_out = nullptr;   // FP64 staging for FP32 Krylov
    std::vector<void*> allocs;
};

// greedy colouring (given or breadth-first vertex order, whichever needs fewer
// colours): colour k holds cells[cptr[k] .. cptr[k+1])
void color_graph(int n, const std::vector<int>& ptr, const std::vector<int>& adj, std::vector<int>& cptr, std::vector<int>& cells,
                 int& ncol);
// build the hierarchy for the interior face graph of the mesh (host, once);
// allocations are recorded in mg.allocs
int mg_build(const DevGrid& g, const std::vector<int>& par, const std::vector<int>& nei, Multigrid& mg, int max_levels,
             int coarse_cells);
void mg_free(Multigrid& mg);
// re-assemble every coarse level from the fine FD blocks and invert the
// diagonal blocks of all levels (fine_Dinv: FP64 fine-level inverse)
void mg_setup(const DevGrid& g, Multigrid& mg, const double* D, const double* JL, const double* JR, double* fine_Dinv, int* status,
              hipStream_t s);
// x = M^{-1} b by one V-cycle (b, x: [nv][ld] FP64 fine vectors); trans: the
// V-cycle of the transposed operator on the same hierarchy (every level's
// blocks transposed: the Galerkin coarse operators of A^T are the transposes
// of A's, since restriction and prolongation are each other's transposes)
void mg_vcycle(const DevGrid& g, const Colors& fine_colors, Multigrid& mg, const double* D, const double* JL, const double* JR,
               const double* fine_Dinv, const double* b, double* x, int sweeps, hipStream_t s, bool trans = false);

}  // namespace fcfd
