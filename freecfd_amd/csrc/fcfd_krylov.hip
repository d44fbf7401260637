// Linear algebra of the implicit step on the GPU: operator matvec with the
// per-face Jacobian blocks, batched NV x NV block-Jacobi (5x5, or 6x6 with
// the SA model), deterministic reductions and a restarted right-preconditioned
// GMRES, in FP64 or FP32.
//
// Replaces the reference's PETSc MPIAIJ matrix + FGMRES(300) KSP
// (ns_petsc_functions.cc:25-110).  The operator is the reference's implicit
// operator A = P V/dt - dR/dW stored as: diagonal blocks D ([NV*NV][ld]) and
// per face JL, JR ([NV*NV][F]); row P, col N block = -JR^T, row N, col P =
// +JL^T.  Vectors are [NV][ld].
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fcfd_krylov.h"

namespace fcfd {

constexpr int kRedBlocks = 512;   // fixed -> deterministic reductions
constexpr int kRedThreads = 256;

// ---------------------------------------------------------------------------
// element-wise kernels on [NV][ld] vectors (only the first N entries of each
// variable row are meaningful; the loops cover NV*ld for simplicity)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void axpy_kernel(size_t n, T a, const T* __restrict__ x, T* __restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] += a * x[i];
}
template <typename T>
__global__ void scale_copy_kernel(size_t n, T a, const T* __restrict__ x, T* __restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = a * x[i];
}
template <typename A, typename B>
__global__ void convert_kernel(size_t n, const A* __restrict__ x, B* __restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = (B)x[i];
}
template <typename T>
__global__ void xmy_kernel(size_t n, const T* __restrict__ b, T* __restrict__ y) {   // y = b - y
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = b[i] - y[i];
}
// Gram-Schmidt coefficients handed to the update kernels by value (the
// kernel-argument segment), so that the host's coefficients need no copy to
// the device and no wait for it; more than kArgCoefs of them go through
// device memory
constexpr int kArgCoefs = 256;
struct ArgCoefs {
    double c[kArgCoefs];
};
__device__ __forceinline__ double coef(const double* h, int q) { return h[q]; }
__device__ __forceinline__ double coef(const ArgCoefs& h, int q) { return h.c[q]; }

// y = (y - sum_k h[k] * V_k) * scale  (V_k at V + k*stride)
template <typename T, typename CS>
__global__ void mgs_update_kernel(size_t n, int k, const T* __restrict__ V, size_t stride, CS h, T scale, T* __restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        T acc = y[i];
        for (int q = 0; q < k; ++q) acc -= (T)coef(h, q) * V[q * stride + i];
        y[i] = acc * scale;
    }
}
// x += sum_k c[k] * V_k
template <typename T>
__global__ void lincomb_kernel(size_t n, int k, const T* __restrict__ V, size_t stride, const double* __restrict__ c,
                               T* __restrict__ x) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        T acc = x[i];
        for (int q = 0; q < k; ++q) acc += (T)c[q] * V[q * stride + i];
        x[i] = acc;
    }
}

// The same two updates on 16-byte packs (4 floats per lane), the
// basis vectors' packs loaded four at a time before they are used: one
// element at a time, a thread waited on each basis load before it issued the
// next.  Each element is still updated by the same operations in the same
// order (q ascending), so the results are bit-identical.  Needs n, stride
// and the pointers on 16-byte pack boundaries (the launchers check).
template <typename T>
struct alignas(16) Pack16 {
    static constexpr int L = 16 / sizeof(T);
    T v[L];
};
template <typename T, bool UPDATE, typename CS>
__global__ void __launch_bounds__(256) gs_pack_kernel(size_t np, int k, const Pack16<T>* __restrict__ V, size_t pstride, CS h,
                                                      T scale, Pack16<T>* __restrict__ y) {
    constexpr int L = Pack16<T>::L;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < np; i += (size_t)gridDim.x * blockDim.x) {
        Pack16<T> acc = y[i];
        int q = 0;
        for (; q + 4 <= k; q += 4) {
            Pack16<T> a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = V[(size_t)(q + u) * pstride + i];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T c = (T)coef(h, q + u);
#pragma unroll
                for (int l = 0; l < L; ++l) {
                    if (UPDATE) acc.v[l] -= c * a[u].v[l];
                    else acc.v[l] += c * a[u].v[l];
                }
            }
        }
        for (; q < k; ++q) {
            const Pack16<T> a = V[(size_t)q * pstride + i];
            const T c = (T)coef(h, q);
#pragma unroll
            for (int l = 0; l < L; ++l) {
                if (UPDATE) acc.v[l] -= c * a.v[l];
                else acc.v[l] += c * a.v[l];
            }
        }
        if (UPDATE) {
#pragma unroll
            for (int l = 0; l < L; ++l) acc.v[l] = acc.v[l] * scale;
        }
        y[i] = acc;
    }
}

// ---------------------------------------------------------------------------
// deterministic multi-dot: out[q] = <V_q, w> for q < k over interior cells,
// accumulated in FP64; partials [k][kRedBlocks] then a fixed-order tree.
// Each block row takes kDotGroup of the V_q, so w is read once per group
// rather than once per vector; every partial keeps its own order.
// ---------------------------------------------------------------------------
constexpr int kDotGroup = 8;
template <typename T, int NV>
__global__ void __launch_bounds__(kRedThreads) mdot_partial_kernel(int N, size_t ld, int k, const T* __restrict__ V, size_t stride,
                                                                   const T* __restrict__ w, double* __restrict__ part) {
    __shared__ double sh[kDotGroup][kRedThreads];
    const int q0 = blockIdx.y * kDotGroup, nq = min(kDotGroup, k - q0);
    double acc[kDotGroup];
#pragma unroll
    for (int u = 0; u < kDotGroup; ++u) acc[u] = 0.0;
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < N; i += kRedBlocks * kRedThreads) {
        T wv[NV];
#pragma unroll
        for (int r = 0; r < NV; ++r) wv[r] = w[r * ld + i];
#pragma unroll
        for (int u = 0; u < kDotGroup; ++u) {
            if (u < nq) {
                const T* v = V + (size_t)(q0 + u) * stride;
#pragma unroll
                for (int r = 0; r < NV; ++r) acc[u] += (double)v[r * ld + i] * (double)wv[r];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kDotGroup; ++u) sh[u][threadIdx.x] = acc[u];
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int u = 0; u < kDotGroup; ++u) sh[u][threadIdx.x] += sh[u][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < nq) part[(q0 + threadIdx.x) * kRedBlocks + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ void __launch_bounds__(kRedThreads) reduce_final_kernel(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[kRedThreads];
    const int q = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) acc += part[q * kRedBlocks + i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
}

// largest relative change of the absolute pressure and temperature that the
// update dW would make: max over interior cells of |dp|/(p+Pref), |dT|/(T+Tref)
// (a max, so the result does not depend on the reduction order)
__global__ void __launch_bounds__(kRedThreads) relmax_kernel(int N, size_t ld, double Pref, double Tref, const double* __restrict__ W,
                                                             const double* __restrict__ dW, double* __restrict__ part) {
    __shared__ double sh[kRedThreads];
    double m = 0.0;
    for (int c = blockIdx.x * kRedThreads + threadIdx.x; c < N; c += kRedBlocks * kRedThreads) {
        const double rp = fabs(dW[c]) / fabs(W[c] + Pref), rT = fabs(dW[4 * ld + c]) / fabs(W[4 * ld + c] + Tref);
        m = fmax(m, fmax(rp, rT));   // fmax drops a NaN operand: NaNs are caught by the update's own check
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// The reference's update norms of a Newton update dW over interior cells
// (ns.cc:349-381): sums of dp^2/dt^2, |dV|^2/dt^2, dT^2/dt^2, a NaN/Inf count
// and (SA model) dnu~^2/dt^2 -- partial sums per block, fixed-order tree.  W
// is not touched here: apply_update_kernel adds dW only once the count is 0.
template <int NV>
__global__ void __launch_bounds__(kRedThreads) update_norms_kernel(int N, size_t ld, const double* __restrict__ dW,
                                                                   const double* __restrict__ dt, double* __restrict__ part) {
    __shared__ double sh[5][kRedThreads];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = blockIdx.x * kRedThreads + threadIdx.x; c < N; c += kRedBlocks * kRedThreads) {
        double d[NV];
        bool bad = false;
#pragma unroll
        for (int v = 0; v < NV; ++v) { d[v] = dW[v * ld + c]; bad |= !isfinite(d[v]); }
        if (bad) { acc[3] += 1.0; continue; }
        double dt2 = dt[c] * dt[c];
        acc[0] += d[0] * d[0] / dt2;
        acc[1] += d[1] * d[1] / dt2 + d[2] * d[2] / dt2 + d[3] * d[3] / dt2;
        acc[2] += d[4] * d[4] / dt2;
        if constexpr (NV == 6) acc[4] += d[5] * d[5] / dt2;
    }
    for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; ++q) part[q * kRedBlocks + blockIdx.x] = sh[q][0];
}

// W += dW over interior cells (ns.cc:340-348)
template <int NV>
__global__ void __launch_bounds__(256) apply_update_kernel(int N, size_t ld, double* __restrict__ W, const double* __restrict__ dW) {
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= N) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) W[v * ld + c] += dW[v * ld + c];
}

// ---------------------------------------------------------------------------
// operator: y = D x + sum_faces (-JR^T x_N | +JL^T x_P)
// ---------------------------------------------------------------------------
// acc[j] (+ or, SUB, -)= sg B[i][j] x_o[i] over the cell's half-faces in
// order, B = JR (sg = -1) where the cell is the face's parent and JL (sg = +1)
// where it is the neighbour, o the cell across; faces to a boundary ghost
// carry no column (the ghost coupling is in JL).  TRANS: the same row of the
// transposed operator, acc[j] (+/-)= sg B[j][i] x_o[i] with B = JL (sg = +1)
// on the parent side and JR (sg = -1) on the neighbour side.  The half-face entries are
// loaded in one round; each face's loads use valid (clamped) indices and only
// the accumulation is masked, so the loads of several faces can be in flight
// together.  x: [nv][ld] (ghost rows exist, read but unused).
template <typename T, int NV, bool SUB, bool TRANS = false>
__device__ __forceinline__ void offdiag_accumulate(const DevGrid& g, const T* __restrict__ JL, const T* __restrict__ JR,
                                                   const T* __restrict__ x, int c, T (&acc)[NV]) {
    const size_t ld = g.ld, Fs = g.F;
    const int h0 = g.hf_ptr[c], nh = g.hf_ptr[c + 1] - h0;
    int fs[kMaxCellFaces], nb[kMaxCellFaces];
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) {
        const int h = h0 + min(k, nh - 1);
        fs[k] = g.hf_face[h];
        nb[k] = g.hf_nb[h];
    }
#pragma unroll
    for (int k = 0; k < kMaxCellFaces; ++k) {
        const int f = fs[k] >= 0 ? fs[k] : ~fs[k], o = nb[k];
        T xo[NV], blk[NV * NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) xo[i] = x[i * ld + o];
        load_pairs<NV * NV>((fs[k] >= 0) != TRANS ? JR : JL, Fs, f, blk);
        if (k < nh && o < g.N) {
            const T sg = (fs[k] >= 0) != TRANS ? T(-1) : T(1);
#pragma unroll
            for (int i = 0; i < NV; ++i)
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    const T bij = TRANS ? blk[NV * j + i] : blk[NV * i + j];
                    if (SUB) acc[j] -= sg * bij * xo[i];
                    else acc[j] += sg * bij * xo[i];
                }
        }
    }
}

// y = A x, or with TRANS y = A^T x (every block transposed)
template <typename T, int NV, bool TRANS = false>
__global__ void __launch_bounds__(256) bsr_matvec_kernel(DevGrid g, const T* __restrict__ D, const T* __restrict__ JL,
                                                         const T* __restrict__ JR, const T* __restrict__ x, T* __restrict__ y) {
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    T xc[NV], acc[NV], blk[NV * NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) xc[i] = x[i * ld + c];
    load_pairs<NV * NV>(D, ld, c, blk);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < NV; ++i) s += (TRANS ? blk[NV * i + j] : blk[NV * j + i]) * xc[i];
        acc[j] = s;
    }
    offdiag_accumulate<T, NV, false, TRANS>(g, JL, JR, x, c, acc);
#pragma unroll
    for (int j = 0; j < NV; ++j) y[j * ld + c] = acc[j];
}

// sliced-ELL off-diagonal blocks: one thread per (slot, lane) copies the
// face's block with the row's sign (an exact negation), zeros for a pad
template <typename T, int NV>
__global__ void __launch_bounds__(256) sell_pack_kernel(size_t nent, size_t Fs, const int* __restrict__ sface, const T* __restrict__ JL,
                                                        const T* __restrict__ JR, T* __restrict__ blk) {
    using P2 = typename PairOf<T>::type;
    constexpr int NP = pair_groups(NV * NV);
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const size_t slot = e >> 6, lane = e & 63;
    const int fs = sface[e];
    T b[NV * NV];
    if (fs == kSellPad) {
#pragma unroll
        for (int q = 0; q < NV * NV; ++q) b[q] = T(0);
    } else {
        const int f = fs >= 0 ? fs : ~fs;
        load_pairs<NV * NV>(fs >= 0 ? JR : JL, Fs, f, b);
        const T sg = fs >= 0 ? T(-1) : T(1);
#pragma unroll
        for (int q = 0; q < NV * NV; ++q) b[q] = sg * b[q];
    }
    P2* B2 = reinterpret_cast<P2*>(blk);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        P2 t;
        t.x = b[2 * q];
        t.y = 2 * q + 1 < NV * NV ? b[2 * q + 1] : T(0);
        B2[(slot * NP + q) * 64 + lane] = t;
    }
}

// y = D x + off-diagonal blocks x: per cell the diagonal product, then its
// slots in order (acc[j] += b_ij x_i with b already signed: the products and
// sums of bsr_matvec_kernel, which forms (sign b_ij) x_i)
template <typename T, int NV>
__global__ void __launch_bounds__(256) sell_matvec_kernel(DevGrid g, const int* __restrict__ soff, const int* __restrict__ snb,
                                                          const T* __restrict__ D, const T* __restrict__ blk, const T* __restrict__ x,
                                                          T* __restrict__ y) {
    using P2 = typename PairOf<T>::type;
    constexpr int NP = pair_groups(NV * NV);
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const size_t ld = g.ld;
    const int s = c >> 6, lane = c & 63;
    T xc[NV], acc[NV], b[NV * NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) xc[i] = x[i * ld + c];
    load_pairs<NV * NV>(D, ld, c, b);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        T sum = 0;
#pragma unroll
        for (int i = 0; i < NV; ++i) sum += b[NV * j + i] * xc[i];
        acc[j] = sum;
    }
    const int k0 = soff[s], k1 = soff[s + 1];
    const P2* B2 = reinterpret_cast<const P2*>(blk);
#pragma unroll 2
    for (int k = k0; k < k1; ++k) {
        const int o = snb[(size_t)k * 64 + lane];
        if (o >= 0) {
            T xo[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) xo[i] = x[i * ld + o];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const P2 t = B2[((size_t)k * NP + q) * 64 + lane];
                b[2 * q] = t.x;
                if (2 * q + 1 < NV * NV) b[2 * q + 1] = t.y;
            }
#pragma unroll
            for (int i = 0; i < NV; ++i)
#pragma unroll
                for (int j = 0; j < NV; ++j) acc[j] += b[NV * i + j] * xo[i];
        }
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) y[j * ld + c] = acc[j];
}

// batched NV x NV inverse by Gauss-Jordan with partial pivoting (block-Jacobi);
// PAIRED: blocks in row pairs with stride ld (the mesh level), else each
// block contiguous, entry q of block c at c * block_stride(NV) + q (coarse levels)
template <typename T, int NV, bool PAIRED = true>
__global__ void __launch_bounds__(256) bj_invert_kernel(int N, size_t ld, const double* __restrict__ D, T* __restrict__ Dinv,
                                                        int* __restrict__ status) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    T a[NV][2 * NV];
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            a[i][j] = (T)D[PAIRED ? pidx(ld, NV * i + j, c) : (size_t)c * block_stride(NV) + NV * i + j];
            a[i][NV + j] = (i == j) ? T(1) : T(0);
        }
    bool singular = false;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        int piv = k;
        T best = fabs(a[k][k]);
#pragma unroll
        for (int i = k + 1; i < NV; ++i) if (fabs(a[i][k]) > best) { best = fabs(a[i][k]); piv = i; }
        if (best == T(0)) { singular = true; break; }
        // row swap with compile-time indices only (a runtime row index would
        // put the tableau in scratch memory)
#pragma unroll
        for (int i = k + 1; i < NV; ++i) {
            if (i == piv) {
#pragma unroll
                for (int j = 0; j < 2 * NV; ++j) { T t = a[k][j]; a[k][j] = a[i][j]; a[i][j] = t; }
            }
        }
        T inv = T(1) / a[k][k];
#pragma unroll
        for (int j = 0; j < 2 * NV; ++j) a[k][j] *= inv;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            if (i == k) continue;
            T m = a[i][k];
#pragma unroll
            for (int j = 0; j < 2 * NV; ++j) a[i][j] -= m * a[k][j];
        }
    }
    if (singular) atomicOr(status, 2);
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j)
            Dinv[PAIRED ? pidx(ld, NV * i + j, c) : (size_t)c * block_stride(NV) + NV * i + j] =
                singular ? ((i == j) ? T(1) : T(0)) : a[i][NV + j];
}

// y_c = D_c^{-1} x_c, or with TRANS y_c = D_c^{-T} x_c (the adjoint's
// preconditioner)
template <typename T, int NV, bool TRANS = false>
__global__ void __launch_bounds__(256) bj_apply_kernel(int N, size_t ld, const T* __restrict__ Dinv, const T* __restrict__ x,
                                                       T* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    T xc[NV], blk[NV * NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) xc[i] = x[i * ld + c];
    load_pairs<NV * NV>(Dinv, ld, c, blk);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < NV; ++i) s += blk[TRANS ? NV * i + j : NV * j + i] * xc[i];
        y[j * ld + c] = s;
    }
}

// one colour of a multicolour block Gauss-Seidel sweep: for every cell of the
// colour, x_c = D_c^{-1} (b_c - sum over its faces of the off-diagonal block
// times x of the neighbour), with the latest x of the other colours (cells of
// one colour share no face, so the update is race-free); TRANS: the sweep on
// the transposed operator (transposed off-diagonal blocks, D_c^{-T})
template <typename T, int NV, bool TRANS = false>
__global__ void __launch_bounds__(256) bsgs_color_kernel(DevGrid g, const T* __restrict__ Dinv, const T* __restrict__ JL,
                                                         const T* __restrict__ JR, const T* __restrict__ b, T* __restrict__ x,
                                                         const int* __restrict__ cells, int count) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int c = cells[t];
    const size_t ld = g.ld;
    T acc[NV], blk[NV * NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = b[j * ld + c];
    offdiag_accumulate<T, NV, true, TRANS>(g, JL, JR, x, c, acc);
    load_pairs<NV * NV>(Dinv, ld, c, blk);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < NV; ++i) s += (TRANS ? blk[NV * i + j] : blk[NV * j + i]) * acc[i];
        x[j * ld + c] = s;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline dim3 ew_grid(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)); }

template <typename T> void dev_axpy(size_t n, T a, const T* x, T* y, hipStream_t s) {
    hipLaunchKernelGGL(axpy_kernel<T>, ew_grid(n), dim3(256), 0, s, n, a, x, y);
}
template <typename T> void dev_scale_copy(size_t n, T a, const T* x, T* y, hipStream_t s) {
    hipLaunchKernelGGL(scale_copy_kernel<T>, ew_grid(n), dim3(256), 0, s, n, a, x, y);
}
template <typename A, typename B> void dev_convert(size_t n, const A* x, B* y, hipStream_t s) {
    hipLaunchKernelGGL((convert_kernel<A, B>), ew_grid(n), dim3(256), 0, s, n, x, y);
}
template <typename T> void dev_xmy(size_t n, const T* b, T* y, hipStream_t s) {
    hipLaunchKernelGGL(xmy_kernel<T>, ew_grid(n), dim3(256), 0, s, n, b, y);
}

template void dev_axpy<double>(size_t, double, const double*, double*, hipStream_t);
template void dev_axpy<float>(size_t, float, const float*, float*, hipStream_t);
template void dev_scale_copy<double>(size_t, double, const double*, double*, hipStream_t);
template void dev_scale_copy<float>(size_t, float, const float*, float*, hipStream_t);
template void dev_convert<double, float>(size_t, const double*, float*, hipStream_t);
template void dev_convert<float, double>(size_t, const float*, double*, hipStream_t);
template void dev_convert<double, double>(size_t, const double*, double*, hipStream_t);
template void dev_xmy<double>(size_t, const double*, double*, hipStream_t);
template void dev_xmy<float>(size_t, const float*, float*, hipStream_t);

template <typename T>
hipError_t dev_mdot(const DevGrid& g, int k, const T* V, size_t stride, const T* w, double* part, double* out_dev,
                    double* out_host, hipStream_t s) {
    if (g.nv == 6)
        hipLaunchKernelGGL((mdot_partial_kernel<T, 6>), dim3(kRedBlocks, (k + kDotGroup - 1) / kDotGroup), dim3(kRedThreads), 0, s, g.N, (size_t)g.ld, k, V, stride, w, part);
    else if (g.nv == 2)   // the segregated two-equation model's system
        hipLaunchKernelGGL((mdot_partial_kernel<T, 2>), dim3(kRedBlocks, (k + kDotGroup - 1) / kDotGroup), dim3(kRedThreads), 0, s, g.N, (size_t)g.ld, k, V, stride, w, part);
    else
        hipLaunchKernelGGL((mdot_partial_kernel<T, 5>), dim3(kRedBlocks, (k + kDotGroup - 1) / kDotGroup), dim3(kRedThreads), 0, s, g.N, (size_t)g.ld, k, V, stride, w, part);
    hipLaunchKernelGGL(reduce_final_kernel, dim3(k), dim3(kRedThreads), 0, s, part, out_dev);
    hipError_t e = hipMemcpyAsync(out_host, out_dev, k * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}
template hipError_t dev_mdot<double>(const DevGrid&, int, const double*, size_t, const double*, double*, double*, double*, hipStream_t);
template hipError_t dev_mdot<float>(const DevGrid&, int, const float*, size_t, const float*, double*, double*, double*, hipStream_t);

// (FP32 vectors only: with FP64 vectors the packed kernel measured 4% slower
// than the element-wise one on the 5M-cell plate, 0.885 against 0.849 ms)
template <typename T>
static bool on_packs(size_t n, size_t stride, const T* V, const T* y) {
    constexpr size_t L = Pack16<T>::L;
    return sizeof(T) == 4 && n % L == 0 && stride % L == 0 && reinterpret_cast<uintptr_t>(V) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(y) % 16 == 0;
}
template <typename T, typename CS>
static void mgs_update_cs(const DevGrid& g, int k, const T* V, size_t stride, CS h, T scale, T* y, hipStream_t s) {
    size_t n = (size_t)g.nv * g.ld;
    if (on_packs(n, stride, V, y)) {
        constexpr size_t L = Pack16<T>::L;
        hipLaunchKernelGGL((gs_pack_kernel<T, true, CS>), ew_grid(n / L), dim3(256), 0, s, n / L, k,
                           reinterpret_cast<const Pack16<T>*>(V), stride / L, h, scale, reinterpret_cast<Pack16<T>*>(y));
    } else {
        hipLaunchKernelGGL((mgs_update_kernel<T, CS>), ew_grid(n), dim3(256), 0, s, n, k, V, stride, h, scale, y);
    }
}
// coefficients from the host by value (h_host, k <= kArgCoefs) or from h_dev
template <typename T>
void dev_mgs_update(const DevGrid& g, int k, const T* V, size_t stride, const double* h_dev, const double* h_host, T scale, T* y,
                    hipStream_t s) {
    if (h_host && k <= kArgCoefs) {
        ArgCoefs a{};
        for (int q = 0; q < k; ++q) a.c[q] = h_host[q];
        mgs_update_cs<T, ArgCoefs>(g, k, V, stride, a, scale, y, s);
    } else {
        mgs_update_cs<T, const double*>(g, k, V, stride, h_dev, scale, y, s);
    }
}
template <typename T>
void dev_lincomb(const DevGrid& g, int k, const T* V, size_t stride, const double* c_dev, T* x, hipStream_t s) {
    size_t n = (size_t)g.nv * g.ld;
    if (on_packs(n, stride, V, x)) {
        constexpr size_t L = Pack16<T>::L;
        hipLaunchKernelGGL((gs_pack_kernel<T, false, const double*>), ew_grid(n / L), dim3(256), 0, s, n / L, k,
                           reinterpret_cast<const Pack16<T>*>(V), stride / L, c_dev, T(1), reinterpret_cast<Pack16<T>*>(x));
    } else {
        hipLaunchKernelGGL(lincomb_kernel<T>, ew_grid(n), dim3(256), 0, s, n, k, V, stride, c_dev, x);
    }
}
template void dev_mgs_update<double>(const DevGrid&, int, const double*, size_t, const double*, const double*, double, double*,
                                      hipStream_t);
template void dev_mgs_update<float>(const DevGrid&, int, const float*, size_t, const double*, const double*, float, float*,
                                     hipStream_t);
template void dev_lincomb<double>(const DevGrid&, int, const double*, size_t, const double*, double*, hipStream_t);
template void dev_lincomb<float>(const DevGrid&, int, const float*, size_t, const double*, float*, hipStream_t);

template <typename T>
void dev_bsr_matvec(const DevGrid& g, const T* D, const T* JL, const T* JR, const T* x, T* y, hipStream_t s, bool trans) {
    if (trans) {   // the transposed product (the adjoint's multigrid residual)
        if (g.nv == 6) hipLaunchKernelGGL((bsr_matvec_kernel<T, 6, true>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g, D, JL, JR, x, y);
        else hipLaunchKernelGGL((bsr_matvec_kernel<T, 5, true>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g, D, JL, JR, x, y);
        return;
    }
    if (g.nv == 6) hipLaunchKernelGGL((bsr_matvec_kernel<T, 6>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g, D, JL, JR, x, y);
    else if (g.nv == 2) hipLaunchKernelGGL((bsr_matvec_kernel<T, 2>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g, D, JL, JR, x, y);
    else hipLaunchKernelGGL((bsr_matvec_kernel<T, 5>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g, D, JL, JR, x, y);
}
template void dev_bsr_matvec<double>(const DevGrid&, const double*, const double*, const double*, const double*, double*, hipStream_t,
                                     bool);
template void dev_bsr_matvec<float>(const DevGrid&, const float*, const float*, const float*, const float*, float*, hipStream_t,
                                    bool);

template <typename T>
void dev_sell_pack(const DevGrid& g, const SellBlocks& sb, const T* JL, const T* JR, T* blk, hipStream_t s) {
    const size_t nent = sb.nslot * 64;
    if (nent == 0) return;
    const dim3 grid((unsigned)((nent + 255) / 256)), bk(256);
    if (g.nv == 6) hipLaunchKernelGGL((sell_pack_kernel<T, 6>), grid, bk, 0, s, nent, (size_t)g.F, sb.sface, JL, JR, blk);
    else hipLaunchKernelGGL((sell_pack_kernel<T, 5>), grid, bk, 0, s, nent, (size_t)g.F, sb.sface, JL, JR, blk);
}
template <typename T>
void dev_sell_matvec(const DevGrid& g, const SellBlocks& sb, const T* D, const T* blk, const T* x, T* y, hipStream_t s) {
    const dim3 grid = cell_grid(g.N), bk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL((sell_matvec_kernel<T, 6>), grid, bk, 0, s, g, sb.soff, sb.snb, D, blk, x, y);
    else hipLaunchKernelGGL((sell_matvec_kernel<T, 5>), grid, bk, 0, s, g, sb.soff, sb.snb, D, blk, x, y);
}
template void dev_sell_pack<double>(const DevGrid&, const SellBlocks&, const double*, const double*, double*, hipStream_t);
template void dev_sell_pack<float>(const DevGrid&, const SellBlocks&, const float*, const float*, float*, hipStream_t);
template void dev_sell_matvec<double>(const DevGrid&, const SellBlocks&, const double*, const double*, const double*, double*, hipStream_t);
template void dev_sell_matvec<float>(const DevGrid&, const SellBlocks&, const float*, const float*, const float*, float*, hipStream_t);

template <typename T>
void dev_bj_invert(const DevGrid& g, const double* D, T* Dinv, int* status, hipStream_t s) {
    if (g.nv == 6) hipLaunchKernelGGL((bj_invert_kernel<T, 6>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, D, Dinv, status);
    else if (g.nv == 2) hipLaunchKernelGGL((bj_invert_kernel<T, 2>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, D, Dinv, status);
    else hipLaunchKernelGGL((bj_invert_kernel<T, 5>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, D, Dinv, status);
}
template <typename T>
void dev_bj_apply(const DevGrid& g, const T* Dinv, const T* x, T* y, hipStream_t s) {
    if (g.nv == 6) hipLaunchKernelGGL((bj_apply_kernel<T, 6>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, Dinv, x, y);
    else if (g.nv == 2) hipLaunchKernelGGL((bj_apply_kernel<T, 2>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, Dinv, x, y);
    else hipLaunchKernelGGL((bj_apply_kernel<T, 5>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, Dinv, x, y);
}
void dev_bj_apply_t(const DevGrid& g, const double* Dinv, const double* x, double* y, hipStream_t s) {
    if (g.nv == 6)
        hipLaunchKernelGGL((bj_apply_kernel<double, 6, true>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, Dinv, x, y);
    else
        hipLaunchKernelGGL((bj_apply_kernel<double, 5, true>), cell_grid(g.N), dim3(cell_threads(g.N)), 0, s, g.N, (size_t)g.ld, Dinv, x, y);
}
template void dev_bj_invert<double>(const DevGrid&, const double*, double*, int*, hipStream_t);
template void dev_bj_invert<float>(const DevGrid&, const double*, float*, int*, hipStream_t);
template void dev_bj_apply<double>(const DevGrid&, const double*, const double*, double*, hipStream_t);
template void dev_bj_apply<float>(const DevGrid&, const float*, const float*, float*, hipStream_t);

void dev_block_invert(int nv, int n, size_t ld, const double* D, double* Dinv, int* status, hipStream_t s, bool paired) {
    if (n <= 0) return;
    dim3 grid((n + 255) / 256), blk(256);
    if (paired) {
        if (nv == 6) hipLaunchKernelGGL((bj_invert_kernel<double, 6, true>), grid, blk, 0, s, n, ld, D, Dinv, status);
        else hipLaunchKernelGGL((bj_invert_kernel<double, 5, true>), grid, blk, 0, s, n, ld, D, Dinv, status);
    } else {
        if (nv == 6) hipLaunchKernelGGL((bj_invert_kernel<double, 6, false>), grid, blk, 0, s, n, ld, D, Dinv, status);
        else hipLaunchKernelGGL((bj_invert_kernel<double, 5, false>), grid, blk, 0, s, n, ld, D, Dinv, status);
    }
}

template <typename T>
void dev_bsgs_apply(const DevGrid& g, const Colors& col, const T* Dinv, const T* JL, const T* JR, const T* b, T* x, int sweeps,
                    hipStream_t s, bool zero_init, int dir, bool trans) {
    const size_t n = (size_t)g.nv * g.ld;
    if (zero_init) hipLaunchKernelGGL(scale_copy_kernel<T>, ew_grid(n), dim3(256), 0, s, n, T(0), b, x);   // x = 0
    auto one = [&](int k) {
        const int cnt = col.ptr[k + 1] - col.ptr[k];
        if (cnt <= 0) return;
        const int* cells = col.cells + col.ptr[k];
        const dim3 grid = cell_grid(cnt), blk(cell_threads(cnt));
        if (g.nv == 6 && trans) hipLaunchKernelGGL((bsgs_color_kernel<T, 6, true>), grid, blk, 0, s, g, Dinv, JL, JR, b, x, cells, cnt);
        else if (g.nv == 6) hipLaunchKernelGGL((bsgs_color_kernel<T, 6>), grid, blk, 0, s, g, Dinv, JL, JR, b, x, cells, cnt);
        else if (trans) hipLaunchKernelGGL((bsgs_color_kernel<T, 5, true>), grid, blk, 0, s, g, Dinv, JL, JR, b, x, cells, cnt);
        else hipLaunchKernelGGL((bsgs_color_kernel<T, 5>), grid, blk, 0, s, g, Dinv, JL, JR, b, x, cells, cnt);
    };
    for (int sw = 0; sw < sweeps; ++sw) {
        if (dir != 2)
            for (int k = 0; k < col.n; ++k) one(k);        // forward
        if (dir != 1)
            for (int k = col.n - 1; k >= 0; --k) one(k);   // backward
    }
}
template void dev_bsgs_apply<double>(const DevGrid&, const Colors&, const double*, const double*, const double*, const double*, double*,
                                     int, hipStream_t, bool, int, bool);
template void dev_bsgs_apply<float>(const DevGrid&, const Colors&, const float*, const float*, const float*, const float*, float*, int,
                                    hipStream_t, bool, int, bool);

hipError_t dev_relmax(const DevGrid& g, double Pref, double Tref, const double* W, const double* dW, double* part, double* out_host,
                      hipStream_t s) {
    hipLaunchKernelGGL(relmax_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, g.N, (size_t)g.ld, Pref, Tref, W, dW, part);
    std::vector<double> h(kRedBlocks);
    hipError_t e = hipMemcpyAsync(h.data(), part, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    *out_host = m;
    return hipSuccess;
}

hipError_t dev_update(const DevGrid& g, double* W, const double* dW, const double* dt, double* part, double* out_dev,
                      double out_host[5], hipStream_t s) {
    if (g.nv == 6) hipLaunchKernelGGL(update_norms_kernel<6>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, g.N, (size_t)g.ld, dW, dt, part);
    else hipLaunchKernelGGL(update_norms_kernel<5>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, g.N, (size_t)g.ld, dW, dt, part);
    hipLaunchKernelGGL(reduce_final_kernel, dim3(5), dim3(kRedThreads), 0, s, part, out_dev);
    hipError_t e = hipMemcpyAsync(out_host, out_dev, 5 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (out_host[3] > 0.0) return hipSuccess;   // NaN/Inf in dW: W stays as it was
    const dim3 grid = cell_grid(g.N), blk(cell_threads(g.N));
    if (g.nv == 6) hipLaunchKernelGGL(apply_update_kernel<6>, grid, blk, 0, s, g.N, (size_t)g.ld, W, dW);
    else hipLaunchKernelGGL(apply_update_kernel<5>, grid, blk, 0, s, g.N, (size_t)g.ld, W, dW);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// restarted GMRES, right preconditioned: A M^{-1} u = b, x = M^{-1} u
// ---------------------------------------------------------------------------
template <typename T>
int gmres(const DevGrid& g, const LinOp<T>& A, const LinOp<T>& M, const T* b, T* x, const GmresOpts& o, GmresWork<T>& w,
          GmresStats& st, hipStream_t s) {
    const size_t n = (size_t)g.nv * g.ld;
    const int m = o.restart;
    const size_t stride = n;
    T* V = w.V;       // (m+1) vectors
    T* z = w.z;
    T* r = w.r;
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), gv(m + 1, 0.0), h(m + 1, 0.0), y(m, 0.0);
    double bnorm2 = 0.0;
    if (dev_mdot(g, 1, b, stride, b, w.part, w.out_dev, &bnorm2, s) != hipSuccess) return FCFD_EHIP;
    const double bnorm = std::sqrt(bnorm2);
    // initial guess: Knoll (x0 = M^{-1} b) or zero
    if (o.knoll) M.apply(b, x, s);
    else if (hipMemsetAsync(x, 0, n * sizeof(T), s) != hipSuccess) return FCFD_EHIP;
    int it = 0;
    double rnorm = bnorm, r0norm = -1.0;
    st.converged = 0;
    while (it < o.max_iters) {
        // r = b - A x
        A.apply(x, r, s);
        dev_xmy(n, b, r, s);
        double rr = 0.0;
        if (dev_mdot(g, 1, r, stride, r, w.part, w.out_dev, &rr, s) != hipSuccess) return FCFD_EHIP;
        rnorm = std::sqrt(rr);
        if (r0norm < 0.0) r0norm = rnorm;
        double target = std::max(o.rtol * r0norm, o.abstol);
        if (rnorm <= target || rnorm == 0.0) { st.converged = 1; break; }
        dev_scale_copy(n, (T)(1.0 / rnorm), r, V, s);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = rnorm;
        int j = 0;
        for (; j < m && it < o.max_iters; ++j, ++it) {
            T* vj = V + (size_t)j * stride;
            T* vn = V + (size_t)(j + 1) * stride;
            M.apply(vj, z, s);
            A.apply(z, vn, s);
            // classical Gram-Schmidt with selective reorthogonalisation (DGKS,
            // eta = 1/sqrt(2)): one multi-dot per pass returns <V_q, w> and,
            // as its last row, ||w||^2 (w is V_{j+1}); the norm after the
            // projection is ||w||^2 - sum h_q^2, so the normalisation rides on
            // the update kernel.  A second pass runs only when the projection
            // removed more than half of ||w||^2.
            double hn = 0.0;
            for (int pass = 0; pass < 2; ++pass) {
                if (dev_mdot(g, j + 2, V, stride, vn, w.part, w.out_dev, h.data(), s) != hipSuccess) return FCFD_EHIP;
                const double nn = h[j + 1];
                double hs = 0.0;
                for (int q = 0; q <= j; ++q) hs += h[q] * h[q];
                const double est = nn - hs;
                const bool again = pass == 0 && !(est > 0.5 * nn);
                hn = (!again && est > 0.0) ? std::sqrt(est) : 0.0;
                // up to kArgCoefs coefficients ride in the kernel's arguments
                const bool by_value = j + 1 <= kArgCoefs;
                if (!by_value && (hipMemcpyAsync(w.coef_dev, h.data(), (j + 1) * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
                                  hipStreamSynchronize(s) != hipSuccess))
                    return FCFD_EHIP;
                dev_mgs_update(g, j + 1, V, stride, w.coef_dev, by_value ? h.data() : nullptr, (T)(hn > 0.0 ? 1.0 / hn : 1.0), vn, s);
                for (int q = 0; q <= j; ++q) H[(size_t)q * m + j] += h[q];
                if (!again) break;
            }
            H[(size_t)(j + 1) * m + j] = hn;
            // Givens rotations
            for (int q = 0; q < j; ++q) {
                double a = H[(size_t)q * m + j], bq = H[(size_t)(q + 1) * m + j];
                H[(size_t)q * m + j] = cs[q] * a + sn[q] * bq;
                H[(size_t)(q + 1) * m + j] = -sn[q] * a + cs[q] * bq;
            }
            double a = H[(size_t)j * m + j], bq = H[(size_t)(j + 1) * m + j];
            double den = std::sqrt(a * a + bq * bq);
            cs[j] = den > 0.0 ? a / den : 1.0;
            sn[j] = den > 0.0 ? bq / den : 0.0;
            H[(size_t)j * m + j] = den;
            H[(size_t)(j + 1) * m + j] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            rnorm = std::fabs(gv[j + 1]);
            if (rnorm <= std::max(o.rtol * r0norm, o.abstol)) { ++j; ++it; st.converged = 1; break; }
        }
        // solve H y = g (upper triangular j x j), x += M^{-1} (V y)
        for (int q = j - 1; q >= 0; --q) {
            double acc = gv[q];
            for (int k = q + 1; k < j; ++k) acc -= H[(size_t)q * m + k] * y[k];
            y[q] = H[(size_t)q * m + q] != 0.0 ? acc / H[(size_t)q * m + q] : 0.0;
        }
        if (hipMemcpyAsync(w.coef_dev, y.data(), j * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return FCFD_EHIP;
        if (hipMemsetAsync(r, 0, n * sizeof(T), s) != hipSuccess) return FCFD_EHIP;
        dev_lincomb(g, j, V, stride, w.coef_dev, r, s);
        M.apply(r, z, s);
        dev_axpy(n, (T)1, z, x, s);
        std::fill(H.begin(), H.end(), 0.0);
        if (st.converged) break;
    }
    st.iters = it;
    st.rnorm = rnorm;
    st.r0norm = r0norm < 0 ? bnorm : r0norm;
    st.bnorm = bnorm;
    hipError_t e = hipStreamSynchronize(s);
    return e == hipSuccess ? FCFD_OK : FCFD_EHIP;
}

template int gmres<double>(const DevGrid&, const LinOp<double>&, const LinOp<double>&, const double*, double*, const GmresOpts&,
                           GmresWork<double>&, GmresStats&, hipStream_t);
template int gmres<float>(const DevGrid&, const LinOp<float>&, const LinOp<float>&, const float*, float*, const GmresOpts&,
                          GmresWork<float>&, GmresStats&, hipStream_t);

}  // namespace fcfd
