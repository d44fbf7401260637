// Aggregation multigrid preconditioner (see fcfd_mg.h).
//
// Block conventions (those of the FD Jacobian and bsr_matvec_kernel): the
// operator row of cell c is D_c x_c + sum over its faces of a block times the
// other cell's x, the block being -JR_f^T when c is the face's parent and
// +JL_f^T when c is its neighbour (JL/JR in row pairs, entry nv i + j for row
// j, column i).  Coarse-level blocks are row-major and each block contiguous:
// entry (j, i) of block k at [k B + nv j + i], B = block_stride(nv), so a
// thread that walks a row loads each block with a few 16-byte loads from two
// or three cache lines.  A Galerkin source code is (index << 2) | kind with
// kind 0 = a diagonal block of the finer level, 1 = -JR_f^T (parent row,
// neighbour column), 2 = +JL_f^T (neighbour row, parent column), 3 = an
// off-diagonal block of a coarse finer level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fcfd_mg.h"

namespace fcfd {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void load_block(const double* __restrict__ A, size_t k, double (&b)[NV * NV]) {
    const double2* A2 = reinterpret_cast<const double2*>(A + k * block_stride(NV));
#pragma unroll
    for (int q = 0; q < block_stride(NV) / 2; ++q) {
        const double2 t = A2[q];
        b[2 * q] = t.x;
        if (2 * q + 1 < NV * NV) b[2 * q + 1] = t.y;
    }
}
template <int NV>
__device__ __forceinline__ void store_block(double* __restrict__ A, size_t k, const double (&b)[NV * NV]) {
    double2* A2 = reinterpret_cast<double2*>(A + k * block_stride(NV));
#pragma unroll
    for (int q = 0; q < block_stride(NV) / 2; ++q)
        A2[q] = make_double2(b[2 * q], 2 * q + 1 < NV * NV ? b[2 * q + 1] : 0.0);
}

template <int NV, bool FINE>
__global__ void __launch_bounds__(256) mg_assemble_kernel(int nblk, const int* __restrict__ ptr, const int* __restrict__ src,
                                                          double* __restrict__ out, const double* __restrict__ D0, size_t ld0,
                                                          const double* __restrict__ JL, const double* __restrict__ JR, size_t F,
                                                          const double* __restrict__ Dl, const double* __restrict__ Al) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    double acc[NV * NV];
#pragma unroll
    for (int q = 0; q < NV * NV; ++q) acc[q] = 0.0;
    for (int e = ptr[k]; e < ptr[k + 1]; ++e) {
        const int code = src[e];
        const size_t idx = (size_t)(code >> 2);
        const int kind = code & 3;
        if (FINE) {
            double blk[NV * NV];   // the mesh level's blocks are stored in row pairs
            if (kind == 0) {
                load_pairs<NV * NV>(D0, ld0, idx, blk);
#pragma unroll
                for (int q = 0; q < NV * NV; ++q) acc[q] += blk[q];
            } else if (kind == 1) {
                load_pairs<NV * NV>(JR, F, idx, blk);
#pragma unroll
                for (int j = 0; j < NV; ++j)
#pragma unroll
                    for (int i = 0; i < NV; ++i) acc[NV * j + i] -= blk[NV * i + j];
            } else {
                load_pairs<NV * NV>(JL, F, idx, blk);
#pragma unroll
                for (int j = 0; j < NV; ++j)
#pragma unroll
                    for (int i = 0; i < NV; ++i) acc[NV * j + i] += blk[NV * i + j];
            }
        } else {
            double blk[NV * NV];
            load_block<NV>(kind == 0 ? Dl : Al, idx, blk);
#pragma unroll
            for (int q = 0; q < NV * NV; ++q) acc[q] += blk[q];
        }
    }
    store_block<NV>(out, k, acc);
}

// block k, transposed with TRANS (the transposed operator's block)
template <int NV, bool TRANS>
__device__ __forceinline__ void load_block_t(const double* __restrict__ A, size_t k, double (&b)[NV * NV]) {
    if constexpr (TRANS) {
        double t[NV * NV];
        load_block<NV>(A, k, t);
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
            for (int i = 0; i < NV; ++i) b[NV * j + i] = t[NV * i + j];
    } else {
        load_block<NV>(A, k, b);
    }
}

// block Gauss-Seidel update of coarse row I: x_I = D_I^{-1} (b_I - sum of
// the row's off-diagonal blocks times x).  TRANS: the transposed operator,
// whose row I holds the blocks A(J, I)^T (rptr/col/idx: the transposed
// pattern, idx the block of A each entry transposes) and D_I^{-T}.
template <int NV, bool TRANS = false>
__device__ __forceinline__ void sgs_row(int I, size_t ld, const int* __restrict__ rptr, const int* __restrict__ col,
                                        const int* __restrict__ idx, const double* __restrict__ A, const double* __restrict__ Dinv,
                                        const double* __restrict__ b, double* __restrict__ x) {
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = b[j * ld + I];
    for (int k = rptr[I]; k < rptr[I + 1]; ++k) {
        const int J = col[k];
        double xj[NV], a[NV * NV];
        load_block_t<NV, TRANS>(A, TRANS ? idx[k] : k, a);
#pragma unroll
        for (int i = 0; i < NV; ++i) xj[i] = x[i * ld + J];
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
            for (int i = 0; i < NV; ++i) acc[j] -= a[NV * j + i] * xj[i];
    }
    double di[NV * NV];
    load_block_t<NV, TRANS>(Dinv, I, di);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NV; ++i) s += di[NV * j + i] * acc[i];
        x[j * ld + I] = s;
    }
}

// one colour of a block Gauss-Seidel sweep on a coarse level
template <int NV, bool TRANS = false>
__global__ void __launch_bounds__(256) mg_sgs_color_kernel(const int* __restrict__ cells, int count, size_t ld,
                                                           const int* __restrict__ rptr, const int* __restrict__ col,
                                                           const int* __restrict__ idx, const double* __restrict__ A,
                                                           const double* __restrict__ Dinv, const double* __restrict__ b,
                                                           double* __restrict__ x) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    sgs_row<NV, TRANS>(cells[t], ld, rptr, col, idx, A, Dinv, b, x);
}

// all the colours of `sweeps` sweeps on a level of at most kSmallLevel cells
// in one workgroup, a workgroup barrier between colours; the same updates in
// the same colour order as the per-colour launches (sgs_row's operations), in
// one launch instead of one per colour.  The level's x stays in LDS for the
// whole call (one gather round per colour fewer; the blocks, b and the
// row structure are read from memory).  dir as in coarse_sgs.
constexpr int kSmallLevel = 1024;
template <int NV, bool TRANS = false>
__global__ void __launch_bounds__(kSmallLevel) mg_sgs_level_kernel(int n, int ncol, const int* __restrict__ cptr,
                                                                   const int* __restrict__ cells, size_t ld,
                                                                   const int* __restrict__ rptr, const int* __restrict__ col,
                                                                   const int* __restrict__ idx,
                                                                   const double* __restrict__ A, const double* __restrict__ Dinv,
                                                                   const double* __restrict__ b, double* __restrict__ x, int sweeps,
                                                                   int dir) {
    __shared__ double xs[NV * kSmallLevel];
    for (int i = threadIdx.x; i < n; i += blockDim.x)
#pragma unroll
        for (int v = 0; v < NV; ++v) xs[v * kSmallLevel + i] = x[v * ld + i];
    __syncthreads();
    for (int sw = 0; sw < sweeps; ++sw) {
        for (int pass = 0; pass < 2; ++pass) {
            if ((pass == 0 && dir == 2) || (pass == 1 && dir == 1)) continue;
            for (int q = 0; q < ncol; ++q) {
                const int k = pass == 0 ? q : ncol - 1 - q;
                for (int t = cptr[k] + threadIdx.x; t < cptr[k + 1]; t += blockDim.x) {
                    const int I = cells[t];
                    double acc[NV];
#pragma unroll
                    for (int j = 0; j < NV; ++j) acc[j] = b[j * ld + I];
                    for (int e = rptr[I]; e < rptr[I + 1]; ++e) {
                        const int J = col[e];
                        double xj[NV], a[NV * NV];
                        load_block_t<NV, TRANS>(A, TRANS ? idx[e] : e, a);
#pragma unroll
                        for (int i = 0; i < NV; ++i) xj[i] = xs[i * kSmallLevel + J];
#pragma unroll
                        for (int j = 0; j < NV; ++j)
#pragma unroll
                            for (int i = 0; i < NV; ++i) acc[j] -= a[NV * j + i] * xj[i];
                    }
                    double di[NV * NV];
                    load_block_t<NV, TRANS>(Dinv, I, di);
#pragma unroll
                    for (int j = 0; j < NV; ++j) {
                        double s = 0.0;
#pragma unroll
                        for (int i = 0; i < NV; ++i) s += di[NV * j + i] * acc[i];
                        xs[j * kSmallLevel + I] = s;
                    }
                }
                __syncthreads();
            }
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x)
#pragma unroll
        for (int v = 0; v < NV; ++v) x[v * ld + i] = xs[v * kSmallLevel + i];
}

// r = b - A x on a coarse level
template <int NV, bool TRANS = false>
__global__ void __launch_bounds__(256) mg_residual_kernel(int n, size_t ld, const int* __restrict__ rptr,
                                                          const int* __restrict__ col, const int* __restrict__ idx,
                                                          const double* __restrict__ D, const double* __restrict__ A,
                                                          const double* __restrict__ b, const double* __restrict__ x,
                                                          double* __restrict__ r) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    double acc[NV], xi[NV];
    double d[NV * NV];
    load_block_t<NV, TRANS>(D, I, d);
#pragma unroll
    for (int j = 0; j < NV; ++j) { acc[j] = b[j * ld + I]; xi[j] = x[j * ld + I]; }
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[j] -= d[NV * j + i] * xi[i];
    for (int k = rptr[I]; k < rptr[I + 1]; ++k) {
        const int J = col[k];
        double xj[NV], a[NV * NV];
        load_block_t<NV, TRANS>(A, TRANS ? idx[k] : k, a);
#pragma unroll
        for (int i = 0; i < NV; ++i) xj[i] = x[i * ld + J];
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
            for (int i = 0; i < NV; ++i) acc[j] -= a[NV * j + i] * xj[i];
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) r[j * ld + I] = acc[j];
}

// bc[I] = sum of rf over the members of aggregate I (member order fixed)
template <int NV>
__global__ void __launch_bounds__(256) mg_restrict_kernel(int n, const int* __restrict__ mptr, const int* __restrict__ mem,
                                                          const double* __restrict__ rf, size_t ldf, double* __restrict__ bc,
                                                          size_t ldc) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    double s[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) s[j] = 0.0;
    for (int e = mptr[I]; e < mptr[I + 1]; ++e) {
        const int i = mem[e];
#pragma unroll
        for (int j = 0; j < NV; ++j) s[j] += rf[j * ldf + i];
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) bc[j * ldc + I] = s[j];
}

// xf[i] += xc[agg[i]]
template <int NV>
__global__ void __launch_bounds__(256) mg_prolong_kernel(int nf, const int* __restrict__ agg, const double* __restrict__ xc, size_t ldc,
                                                         double* __restrict__ xf, size_t ldf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int I = agg[i];
#pragma unroll
    for (int j = 0; j < NV; ++j) xf[j * ldf + i] += xc[j * ldc + I];
}

__global__ void __launch_bounds__(256) mg_zero_kernel(size_t n, double* __restrict__ x) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host: hierarchy
// ---------------------------------------------------------------------------
namespace {

template <typename T>
int mg_upload(Multigrid& mg, const std::vector<T>& h, const T** out) {
    void* p = nullptr;
    size_t n = std::max<size_t>(h.size(), 1);
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return FCFD_EHIP;
    mg.allocs.push_back(p);
    if (!h.empty() && hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return FCFD_EHIP;
    *out = (const T*)p;
    return FCFD_OK;
}

int mg_alloc(Multigrid& mg, size_t n, double** out) {
    void* p = nullptr;
    n = std::max<size_t>(n, 1);
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return FCFD_EHIP;
    mg.allocs.push_back(p);
    if (hipMemset(p, 0, n * sizeof(double)) != hipSuccess) return FCFD_EHIP;
    *out = (double*)p;
    return FCFD_OK;
}

// plain aggregation of an undirected graph (CSR, no self loops): a free cell
// whose neighbours are all free seeds an aggregate with them; then every
// leftover cell joins the first-phase aggregate of its first aggregated
// neighbour; isolated leftovers form their own aggregates
int aggregate(int n, const std::vector<int>& ptr, const std::vector<int>& adj, std::vector<int>& agg) {
    agg.assign(n, -1);
    int na = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool free_nb = true;
        for (int k = ptr[i]; k < ptr[i + 1] && free_nb; ++k) free_nb = agg[adj[k]] < 0;
        if (!free_nb) continue;
        agg[i] = na;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) agg[adj[k]] = na;
        ++na;
    }
    std::vector<int> first(agg);
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k)
            if (first[adj[k]] >= 0) { agg[i] = first[adj[k]]; break; }
    }
    for (int i = 0; i < n; ++i)
        if (agg[i] < 0) agg[i] = na++;
    return na;
}

}  // namespace

// Greedy colouring of an undirected graph (CSR, no self loops), tried in the
// given vertex order and in breadth-first order (per component, from its
// lowest vertex); the order with fewer colours is kept (the given order on a
// tie).  Breadth-first greedy colours a bipartite graph -- the face graph of a
// hexahedral or quadrilateral mesh -- with two colours.  Each colour is a
// dependent launch of a Gauss-Seidel sweep, so fewer colours mean fewer
// launches.
void color_graph(int n, const std::vector<int>& ptr, const std::vector<int>& adj, std::vector<int>& cptr, std::vector<int>& cells,
                 int& ncol) {
    auto greedy = [&](const std::vector<int>& order, std::vector<int>& col) {
        col.assign(n, -1);
        int nc = 0;
        std::vector<char> used;
        for (int i : order) {
            used.assign(nc + 1, 0);
            for (int k = ptr[i]; k < ptr[i + 1]; ++k)
                if (col[adj[k]] >= 0) used[col[adj[k]]] = 1;
            int c = 0;
            while (c < nc && used[c]) ++c;
            col[i] = c;
            if (c == nc) ++nc;
        }
        return nc;
    };
    std::vector<int> natural(n), bfs;
    for (int i = 0; i < n; ++i) natural[i] = i;
    bfs.reserve(n);
    {
        std::vector<char> seen(n, 0);
        for (int s0 = 0; s0 < n; ++s0) {
            if (seen[s0]) continue;
            size_t head = bfs.size();
            bfs.push_back(s0);
            seen[s0] = 1;
            while (head < bfs.size()) {
                const int v = bfs[head++];
                for (int k = ptr[v]; k < ptr[v + 1]; ++k)
                    if (!seen[adj[k]]) { seen[adj[k]] = 1; bfs.push_back(adj[k]); }
            }
        }
    }
    std::vector<int> c1, c2;
    const int n1 = greedy(natural, c1), n2 = greedy(bfs, c2);
    const std::vector<int>& col = n2 < n1 ? c2 : c1;
    ncol = n2 < n1 ? n2 : n1;
    cptr.assign(ncol + 1, 0);
    for (int i = 0; i < n; ++i) cptr[col[i] + 1]++;
    for (int c = 0; c < ncol; ++c) cptr[c + 1] += cptr[c];
    std::vector<int> next(cptr.begin(), cptr.end() - 1);
    cells.assign(n, 0);
    for (int i = 0; i < n; ++i) cells[next[col[i]]++] = i;
}

int mg_build(const DevGrid& g, const std::vector<int>& par, const std::vector<int>& nei, Multigrid& mg, int max_levels,
             int coarse_cells) {
    const int N = g.N, F = g.F, nv = g.nv;
    mg.nv = nv;
    mg.lv.clear();
    // level 0 adjacency: interior faces (both cells interior)
    std::vector<int> ptr(N + 1, 0), adj;
    for (int f = 0; f < F; ++f)
        if (nei[f] < N) { ptr[par[f] + 1]++; ptr[nei[f] + 1]++; }
    for (int i = 0; i < N; ++i) ptr[i + 1] += ptr[i];
    adj.resize(ptr[N]);
    {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int f = 0; f < F; ++f)
            if (nei[f] < N) { adj[fill[par[f]]++] = nei[f]; adj[fill[nei[f]]++] = par[f]; }
    }
    MgLevel l0;
    l0.n = N;
    l0.ld = g.ld;
    mg.lv.push_back(l0);
    int rc;
    if ((rc = mg_alloc(mg, (size_t)nv * g.ld, &mg.fine_r)) || (rc = mg_alloc(mg, (size_t)nv * g.ld, &mg.fine_in)) ||
        (rc = mg_alloc(mg, (size_t)nv * g.ld, &mg.fine_out)))
        return rc;
    // the finer level's blocks as (row, col) pairs with their source codes,
    // diagonal sources per row
    std::vector<int> erow, ecol, ecode;   // off-diagonal blocks of the finer level
    for (int f = 0; f < F; ++f) {
        if (nei[f] >= N) continue;
        erow.push_back(par[f]); ecol.push_back(nei[f]); ecode.push_back((f << 2) | 1);
        erow.push_back(nei[f]); ecol.push_back(par[f]); ecode.push_back((f << 2) | 2);
    }
    int nf = N;
    for (int level = 1; level < max_levels && nf > coarse_cells; ++level) {
        std::vector<int> agg;
        const int nc = aggregate(nf, ptr, adj, agg);
        if (nc >= nf || nc < 1) break;
        MgLevel L;
        L.n = nc;
        L.ld = ((nc + 63) / 64) * 64;
        // members
        std::vector<int> mptr(nc + 1, 0), mem(nf);
        for (int i = 0; i < nf; ++i) mptr[agg[i] + 1]++;
        for (int I = 0; I < nc; ++I) mptr[I + 1] += mptr[I];
        {
            std::vector<int> next(mptr.begin(), mptr.end() - 1);
            for (int i = 0; i < nf; ++i) mem[next[agg[i]]++] = i;
        }
        // coarse adjacency (sorted rows) from the finer off-diagonal blocks
        std::vector<std::vector<int>> rows(nc);
        for (size_t e = 0; e < erow.size(); ++e) {
            int I = agg[erow[e]], J = agg[ecol[e]];
            if (I != J) rows[I].push_back(J);
        }
        std::vector<int> rptr(nc + 1, 0), col;
        for (int I = 0; I < nc; ++I) {
            auto& r = rows[I];
            std::sort(r.begin(), r.end());
            r.erase(std::unique(r.begin(), r.end()), r.end());
            rptr[I + 1] = rptr[I] + (int)r.size();
            col.insert(col.end(), r.begin(), r.end());
        }
        L.nnz = (int)col.size();
        // Galerkin source lists: diagonal = member diagonals, then internal
        // off-diagonal blocks (finer-level order); off-diagonal (I,J) = the
        // finer blocks from I to J (finer-level order)
        const int kdiag = 0, koff = level == 1 ? -1 : 3;
        std::vector<std::vector<int>> dl(nc), ol(L.nnz);
        for (int I = 0; I < nc; ++I)
            for (int e = mptr[I]; e < mptr[I + 1]; ++e) dl[I].push_back((mem[e] << 2) | kdiag);
        for (size_t e = 0; e < erow.size(); ++e) {
            const int I = agg[erow[e]], J = agg[ecol[e]];
            const int code = koff < 0 ? ecode[e] : (((int)e) << 2) | koff;
            if (I == J) {
                dl[I].push_back(code);
            } else {
                const int k = (int)(std::lower_bound(col.begin() + rptr[I], col.begin() + rptr[I + 1], J) - col.begin());
                ol[k].push_back(code);
            }
        }
        std::vector<int> dptr(nc + 1, 0), dsrc, optr(L.nnz + 1, 0), osrc;
        for (int I = 0; I < nc; ++I) { dptr[I + 1] = dptr[I] + (int)dl[I].size(); dsrc.insert(dsrc.end(), dl[I].begin(), dl[I].end()); }
        for (int k = 0; k < L.nnz; ++k) { optr[k + 1] = optr[k] + (int)ol[k].size(); osrc.insert(osrc.end(), ol[k].begin(), ol[k].end()); }
        // colours
        std::vector<int> cptr, cells;
        int ncol = 0;
        color_graph(nc, rptr, col, cptr, cells, ncol);
        L.rptr_h = rptr;
        // transposed pattern: entry (J, k) of row I for each block k = A(J, I),
        // rows filled in ascending J
        std::vector<int> tptr(nc + 1, 0), tcol(L.nnz), tidx(L.nnz);
        for (int k = 0; k < L.nnz; ++k) tptr[col[k] + 1]++;
        for (int I = 0; I < nc; ++I) tptr[I + 1] += tptr[I];
        {
            std::vector<int> next(tptr.begin(), tptr.end() - 1);
            for (int J = 0; J < nc; ++J)
                for (int k = rptr[J]; k < rptr[J + 1]; ++k) {
                    const int q = next[col[k]]++;
                    tcol[q] = J;
                    tidx[q] = k;
                }
        }
        if ((rc = mg_upload(mg, tptr, &L.tptr)) || (rc = mg_upload(mg, tcol, &L.tcol)) || (rc = mg_upload(mg, tidx, &L.tidx))) return rc;
        if ((rc = mg_upload(mg, agg, &L.agg)) || (rc = mg_upload(mg, mptr, &L.mptr)) || (rc = mg_upload(mg, mem, &L.mem)) ||
            (rc = mg_upload(mg, rptr, &L.rptr)) || (rc = mg_upload(mg, col, &L.col)) || (rc = mg_upload(mg, dptr, &L.dptr)) ||
            (rc = mg_upload(mg, dsrc, &L.dsrc)) || (rc = mg_upload(mg, optr, &L.optr)) || (rc = mg_upload(mg, osrc, &L.osrc)) ||
            (rc = mg_upload(mg, cells, &L.colors.cells)) || (rc = mg_upload(mg, cptr, &L.cptr_dev)))
            return rc;
        L.colors.n = ncol;
        L.colors.ptr = cptr;
        // FCFD_MG_NO_LEVEL_LAUNCH=1: one launch per colour on every level (the
        // same updates; for checking the one-launch sweeps against them)
        const char* e = std::getenv("FCFD_MG_NO_LEVEL_LAUNCH");
        L.one_launch = nc <= kSmallLevel && !(e && std::strcmp(e, "1") == 0);
        const size_t b2 = (size_t)block_stride(nv);   // entries per stored block (nv^2 and a pad entry when odd)
        if ((rc = mg_alloc(mg, b2 * nc, &L.D)) || (rc = mg_alloc(mg, b2 * std::max(L.nnz, 1), &L.A)) ||
            (rc = mg_alloc(mg, b2 * nc, &L.Dinv)) || (rc = mg_alloc(mg, (size_t)nv * L.ld, &L.b)) ||
            (rc = mg_alloc(mg, (size_t)nv * L.ld, &L.x)) || (rc = mg_alloc(mg, (size_t)nv * L.ld, &L.r)))
            return rc;
        mg.lv.push_back(L);
        // next level's graph and off-diagonal block list
        ptr = rptr;
        adj = col;
        erow.assign(L.nnz, 0);
        ecol = col;
        ecode.assign(L.nnz, 0);
        for (int I = 0; I < nc; ++I)
            for (int k = rptr[I]; k < rptr[I + 1]; ++k) erow[k] = I;
        nf = nc;
    }
    mg.ready = mg.lv.size() > 1;
    return FCFD_OK;
}

void mg_free(Multigrid& mg) {
    for (void* p : mg.allocs) (void)hipFree(p);
    mg.allocs.clear();
    mg.lv.clear();
    mg.ready = false;
}

// ---------------------------------------------------------------------------
// host: set-up and V-cycle
// ---------------------------------------------------------------------------
template <int NV>
static void setup_t(const DevGrid& g, Multigrid& mg, const double* D, const double* JL, const double* JR, int* status, hipStream_t s) {
    for (size_t l = 1; l < mg.lv.size(); ++l) {
        MgLevel& L = mg.lv[l];
        if (l == 1) {
            hipLaunchKernelGGL((mg_assemble_kernel<NV, true>), dim3((L.n + 255) / 256), dim3(256), 0, s, L.n, L.dptr, L.dsrc, L.D, D,
                               (size_t)g.ld, JL, JR, (size_t)g.F, nullptr, nullptr);
            if (L.nnz > 0)
                hipLaunchKernelGGL((mg_assemble_kernel<NV, true>), dim3((L.nnz + 255) / 256), dim3(256), 0, s, L.nnz, L.optr, L.osrc, L.A,
                                   D, (size_t)g.ld, JL, JR, (size_t)g.F, nullptr, nullptr);
        } else {
            const MgLevel& P = mg.lv[l - 1];
            hipLaunchKernelGGL((mg_assemble_kernel<NV, false>), dim3((L.n + 255) / 256), dim3(256), 0, s, L.n, L.dptr, L.dsrc, L.D,
                               nullptr, (size_t)0, nullptr, nullptr, (size_t)0, P.D, P.A);
            if (L.nnz > 0)
                hipLaunchKernelGGL((mg_assemble_kernel<NV, false>), dim3((L.nnz + 255) / 256), dim3(256), 0, s, L.nnz, L.optr, L.osrc,
                                   L.A, nullptr, (size_t)0, nullptr, nullptr, (size_t)0, P.D, P.A);
        }
        dev_block_invert(NV, L.n, (size_t)0, L.D, L.Dinv, status, s, false);
    }
}

void mg_setup(const DevGrid& g, Multigrid& mg, const double* D, const double* JL, const double* JR, double* fine_Dinv, int* status,
              hipStream_t s) {
    dev_block_invert(g.nv, g.N, (size_t)g.ld, D, fine_Dinv, status, s, true);
    if (g.nv == 6) setup_t<6>(g, mg, D, JL, JR, status, s);
    else setup_t<5>(g, mg, D, JL, JR, status, s);
}

template <int NV, bool TRANS>
static void coarse_sgs(const MgLevel& L, int sweeps, hipStream_t s, int dir) {
    const int* rp = TRANS ? L.tptr : L.rptr;
    const int* cl = TRANS ? L.tcol : L.col;
    const int* ix = TRANS ? L.tidx : nullptr;
    if (L.one_launch) {   // small level: every colour in one workgroup launch
        hipLaunchKernelGGL((mg_sgs_level_kernel<NV, TRANS>), dim3(1), dim3(kSmallLevel), 0, s, L.n, L.colors.n, L.cptr_dev,
                           L.colors.cells, (size_t)L.ld, rp, cl, ix, L.A, L.Dinv, L.b, L.x, sweeps, dir);
        return;
    }
    auto one = [&](int c) {
        const int cnt = L.colors.ptr[c + 1] - L.colors.ptr[c];
        if (cnt <= 0) return;
        hipLaunchKernelGGL((mg_sgs_color_kernel<NV, TRANS>), cell_grid(cnt), dim3(cell_threads(cnt)), 0, s,
                           L.colors.cells + L.colors.ptr[c], cnt, (size_t)L.ld, rp, cl, ix, L.A, L.Dinv, L.b, L.x);
    };
    for (int sw = 0; sw < sweeps; ++sw) {
        if (dir != 2)
            for (int c = 0; c < L.colors.n; ++c) one(c);
        if (dir != 1)
            for (int c = L.colors.n - 1; c >= 0; --c) one(c);
    }
}

template <int NV, bool TRANS>
static void level_cycle(Multigrid& mg, size_t l, int sweeps, hipStream_t s) {
    MgLevel& L = mg.lv[l];
    const size_t nvec = (size_t)NV * L.ld;
    hipLaunchKernelGGL(mg_zero_kernel, dim3((unsigned)std::min<size_t>((nvec + 255) / 256, 4096)), dim3(256), 0, s, nvec, L.x);
    if (l + 1 == mg.lv.size()) {   // coarsest: smoothing only
        coarse_sgs<NV, TRANS>(L, 2 * sweeps, s, 0);
        return;
    }
    coarse_sgs<NV, TRANS>(L, sweeps, s, 1);   // pre-smoothing: forward sweeps
    hipLaunchKernelGGL((mg_residual_kernel<NV, TRANS>), cell_grid(L.n), dim3(cell_threads(L.n)), 0, s, L.n, (size_t)L.ld,
                       TRANS ? L.tptr : L.rptr, TRANS ? L.tcol : L.col, TRANS ? L.tidx : nullptr, L.D, L.A, L.b, L.x, L.r);
    MgLevel& C = mg.lv[l + 1];
    hipLaunchKernelGGL(mg_restrict_kernel<NV>, cell_grid(C.n), dim3(cell_threads(C.n)), 0, s, C.n, C.mptr, C.mem, L.r, (size_t)L.ld, C.b,
                       (size_t)C.ld);
    level_cycle<NV, TRANS>(mg, l + 1, sweeps, s);
    hipLaunchKernelGGL(mg_prolong_kernel<NV>, cell_grid(L.n), dim3(cell_threads(L.n)), 0, s, L.n, C.agg, C.x, (size_t)C.ld, L.x,
                       (size_t)L.ld);
    coarse_sgs<NV, TRANS>(L, sweeps, s, 2);   // post-smoothing: backward sweeps
}

template <int NV, bool TRANS>
static void vcycle_t(const DevGrid& g, const Colors& fc, Multigrid& mg, const double* D, const double* JL, const double* JR,
                     const double* Dinv, const double* b, double* x, int sweeps, hipStream_t s) {
    const size_t n5 = (size_t)NV * g.ld;
    dev_bsgs_apply<double>(g, fc, Dinv, JL, JR, b, x, sweeps, s, true, 1, TRANS);   // forward pre-smoothing
    dev_bsr_matvec<double>(g, D, JL, JR, x, mg.fine_r, s, TRANS);
    dev_xmy<double>(n5, b, mg.fine_r, s);   // r = b - A x
    MgLevel& C = mg.lv[1];
    hipLaunchKernelGGL(mg_restrict_kernel<NV>, cell_grid(C.n), dim3(cell_threads(C.n)), 0, s, C.n, C.mptr, C.mem, mg.fine_r,
                       (size_t)g.ld, C.b, (size_t)C.ld);
    level_cycle<NV, TRANS>(mg, 1, sweeps, s);
    hipLaunchKernelGGL(mg_prolong_kernel<NV>, cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, C.agg, C.x, (size_t)C.ld, x,
                       (size_t)g.ld);
    dev_bsgs_apply<double>(g, fc, Dinv, JL, JR, b, x, sweeps, s, false, 2, TRANS);   // backward post-smoothing
}

void mg_vcycle(const DevGrid& g, const Colors& fine_colors, Multigrid& mg, const double* D, const double* JL, const double* JR,
               const double* fine_Dinv, const double* b, double* x, int sweeps, hipStream_t s, bool trans) {
    if (g.nv == 6 && trans) vcycle_t<6, true>(g, fine_colors, mg, D, JL, JR, fine_Dinv, b, x, sweeps, s);
    else if (g.nv == 6) vcycle_t<6, false>(g, fine_colors, mg, D, JL, JR, fine_Dinv, b, x, sweeps, s);
    else if (trans) vcycle_t<5, true>(g, fine_colors, mg, D, JL, JR, fine_Dinv, b, x, sweeps, s);
    else vcycle_t<5, false>(g, fine_colors, mg, D, JL, JR, fine_Dinv, b, x, sweeps, s);
}

}  // namespace fcfd
