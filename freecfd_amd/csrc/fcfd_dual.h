// Forward-mode dual numbers for the tangent (JVP) kernels.
//
// The value part of every operation is the same IEEE operation the FP64
// residual performs, so a tangent kernel reproduces the residual bitwise in
// .v while .d carries the directional derivative.  Branches take the side
// selected by the value (the active branch of min/max/|x| and of the flux
// switches).  The derivative lane has its own type TD: double (Dual, the
// exact tangent of the FP64 residual) or float (DualF: the FP32 tangent of
// the FP32 Krylov solver).  Sums and products of derivatives are TD
// arithmetic (the values they multiply by, rounded to TD, are the flow's
// states and geometry, well inside FP32's range); a quotient, square root or
// power forms its derivative in FP64 and rounds it once, since a ratio or a
// divisor can leave FP32's range on its own (a limiter ratio over a
// vanishing slope).  With TD
// = double every conversion is the identity and the expressions are those of
// a plain double dual number.
#pragma once
#include <hip/hip_runtime.h>

namespace fcfd {

template <typename TD>
struct DualT {
    double v;
    TD d;
    __host__ __device__ DualT() : v(0.0), d(TD(0)) {}
    __host__ __device__ DualT(double a) : v(a), d(TD(0)) {}
    __host__ __device__ DualT(double a, TD b) : v(a), d(b) {}
};
using Dual = DualT<double>;
using DualF = DualT<float>;

template <typename TD> __device__ __forceinline__ double val(const DualT<TD>& x) { return x.v; }

template <typename TD> __device__ __forceinline__ DualT<TD> operator+(const DualT<TD>& a, const DualT<TD>& b) { return DualT<TD>(a.v + b.v, a.d + b.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator+(const DualT<TD>& a, double b) { return DualT<TD>(a.v + b, a.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator+(double a, const DualT<TD>& b) { return DualT<TD>(a + b.v, b.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator-(const DualT<TD>& a, const DualT<TD>& b) { return DualT<TD>(a.v - b.v, a.d - b.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator-(const DualT<TD>& a, double b) { return DualT<TD>(a.v - b, a.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator-(double a, const DualT<TD>& b) { return DualT<TD>(a - b.v, -b.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator-(const DualT<TD>& a) { return DualT<TD>(-a.v, -a.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator*(const DualT<TD>& a, const DualT<TD>& b) {
    return DualT<TD>(a.v * b.v, a.d * TD(b.v) + TD(a.v) * b.d);
}
template <typename TD> __device__ __forceinline__ DualT<TD> operator*(const DualT<TD>& a, double b) { return DualT<TD>(a.v * b, a.d * TD(b)); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator*(double a, const DualT<TD>& b) { return DualT<TD>(a * b.v, TD(a) * b.d); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator/(const DualT<TD>& a, const DualT<TD>& b) {
    double q = a.v / b.v;
    return DualT<TD>(q, TD((double(a.d) - q * double(b.d)) / b.v));
}
template <typename TD> __device__ __forceinline__ DualT<TD> operator/(const DualT<TD>& a, double b) { return DualT<TD>(a.v / b, TD(double(a.d) / b)); }
template <typename TD> __device__ __forceinline__ DualT<TD> operator/(double a, const DualT<TD>& b) {
    double q = a / b.v;
    return DualT<TD>(q, TD(-q * double(b.d) / b.v));
}
template <typename TD> __device__ __forceinline__ DualT<TD>& operator+=(DualT<TD>& a, const DualT<TD>& b) { a = a + b; return a; }
template <typename TD> __device__ __forceinline__ DualT<TD>& operator-=(DualT<TD>& a, const DualT<TD>& b) { a = a - b; return a; }
template <typename TD> __device__ __forceinline__ DualT<TD>& operator*=(DualT<TD>& a, const DualT<TD>& b) { a = a * b; return a; }

// At x = 0 the derivative of sqrt is unbounded; the code takes sqrt(0) only
// for a magnitude |V| of a vector that is exactly zero (the no-slip wall face
// of Steger-Warming), where x = V.V has a zero tangent: the tangent 0 is then
// the exact derivative of |V|^2 and the zero subgradient of |V|.
template <typename TD> __device__ __forceinline__ DualT<TD> sq_root(const DualT<TD>& x) {
    double s = sqrt(x.v);
    return DualT<TD>(s, s > 0.0 ? TD(double(x.d) / (2.0 * s)) : TD(0));
}
template <typename TD> __device__ __forceinline__ DualT<TD> powr(const DualT<TD>& x, double y) {   // y constant
    double p = pow(x.v, y);
    return DualT<TD>(p, TD(y * pow(x.v, y - 1.0) * double(x.d)));
}
template <typename TD> __device__ __forceinline__ DualT<TD> absv(const DualT<TD>& x) { return DualT<TD>(fabs(x.v), x.v < 0.0 ? -x.d : x.d); }

}  // namespace fcfd
