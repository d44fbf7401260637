// Krylov solver and implicit-operator building blocks (fcfd_krylov.hip).
#pragma once
#include <functional>
#include <vector>

#include "fcfd_device.h"

namespace fcfd {

template <typename T>
struct LinOp {   // y = Op(x) on [nv][ld] device vectors, stream-ordered
    std::function<void(const T*, T*, hipStream_t)> apply;
};

struct GmresOpts {
    int restart = 30, max_iters = 30, knoll = 1;
    double rtol = 1e-6, abstol = 1e-12;
};
struct GmresStats {
    int iters = 0, converged = 0;
    double rnorm = 0.0, r0norm = 0.0, bnorm = 0.0;
};
template <typename T>
struct GmresWork {
    T *V = nullptr, *z = nullptr, *r = nullptr;   // (restart+1) x nv ld, nv ld, nv ld
    double *part = nullptr, *out_dev = nullptr, *coef_dev = nullptr;
};

template <typename T> void dev_axpy(size_t n, T a, const T* x, T* y, hipStream_t s);
template <typename T> void dev_scale_copy(size_t n, T a, const T* x, T* y, hipStream_t s);
template <typename A, typename B> void dev_convert(size_t n, const A* x, B* y, hipStream_t s);
template <typename T> void dev_xmy(size_t n, const T* b, T* y, hipStream_t s);
template <typename T>
hipError_t dev_mdot(const DevGrid& g, int k, const T* V, size_t stride, const T* w, double* part, double* out_dev,
                    double* out_host, hipStream_t s);
template <typename T> void dev_bsr_matvec(const DevGrid& g, const T* D, const T* JL, const T* JR, const T* x, T* y, hipStream_t s,
                                          bool trans = false);
// The off-diagonal blocks of the FD Jacobian re-laid for streaming (sliced
// ELL, 64-cell slices as the wavefront): the entries of slice s are slots
// [soff[s], soff[s+1]), 64 lanes each; slot k, lane l holds cell 64 s + l's
// k-th half-face whose other cell is interior (reference face order, the
// order of the matvec's sum), as the signed face id (hf_face's convention)
// in sface and that cell in snb (-1: a pad).  The blocks follow in row
// pairs, pair q of slot k at ((k * pairs + q) * 64 + l), already carrying the
// row's sign (-JR where the cell is the face's parent, +JL where it is the
// neighbour).
constexpr int kSellPad = -2147483647 - 1;   // sface of a pad slot (no face id reaches it)
struct SellBlocks {
    int nslice = 0;
    size_t nslot = 0;
    const int *soff = nullptr, *sface = nullptr, *snb = nullptr;
};
// blk = the signed off-diagonal blocks of JL/JR ([pairs][F] row pairs) in the sliced layout
template <typename T> void dev_sell_pack(const DevGrid& g, const SellBlocks& sb, const T* JL, const T* JR, T* blk, hipStream_t s);
// y = D x + sum of the off-diagonal blocks times x: the same sums in the same
// order as dev_bsr_matvec, bit for bit, with the blocks streamed
template <typename T> void dev_sell_matvec(const DevGrid& g, const SellBlocks& sb, const T* D, const T* blk, const T* x, T* y, hipStream_t s);
template <typename T> void dev_bj_invert(const DevGrid& g, const double* D, T* Dinv, int* status, hipStream_t s);
template <typename T> void dev_bj_apply(const DevGrid& g, const T* Dinv, const T* x, T* y, hipStream_t s);
void dev_bj_apply_t(const DevGrid& g, const double* Dinv, const double* x, double* y, hipStream_t s);   // y = D^{-T} x
// cell colouring of the face graph: colour k holds cells[ptr[k] .. ptr[k+1]) (device), ptr on the host
struct Colors {
    int n = 0;
    std::vector<int> ptr;
    const int* cells = nullptr;
};
// x = M^{-1} b by `sweeps` multicolour block Gauss-Seidel sweeps from x = 0
// (zero_init) or from the x given; dir 0: symmetric (colours forward then
// backward), 1: forward only, 2: backward only; trans: sweeps on the
// transposed operator (the adjoint's preconditioner)
template <typename T>
void dev_bsgs_apply(const DevGrid& g, const Colors& col, const T* Dinv, const T* JL, const T* JR, const T* b, T* x, int sweeps,
                    hipStream_t s, bool zero_init = true, int dir = 0, bool trans = false);
// batched nv x nv inverses of n row-major blocks, stored in row pairs
// (paired, the mesh level's layout, stride ld) or each block contiguous,
// block_stride(nv) entries apart (coarse levels; ld unused)
void dev_block_invert(int nv, int n, size_t ld, const double* D, double* Dinv, int* status, hipStream_t s, bool paired);
// max over interior cells of |dp|/(p+Pref) and |dT|/(T+Tref) for the update dW
hipError_t dev_relmax(const DevGrid& g, double Pref, double Tref, const double* W, const double* dW, double* part, double* out_host,
                      hipStream_t s);
// out_host: sum dp^2/dt^2, sum |dV|^2/dt^2, sum dT^2/dt^2, NaN/Inf count, sum dnu~^2/dt^2;
// W += dW (interior cells) only when the NaN/Inf count is 0, else W is left as it was
hipError_t dev_update(const DevGrid& g, double* W, const double* dW, const double* dt, double* part, double* out_dev,
                      double out_host[5], hipStream_t s);
template <typename T>
int gmres(const DevGrid& g, const LinOp<T>& A, const LinOp<T>& M, const T* b, T* x, const GmresOpts& o, GmresWork<T>& w,
          GmresStats& st, hipStream_t s);

constexpr int kRedPartials = 512 * 8;   // size of GmresWork::part per multi-dot row budget (k <= 8 rows at once)

}  // namespace fcfd
