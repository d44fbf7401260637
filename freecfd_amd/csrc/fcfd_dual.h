// Forward-mode dual numbers for the tangent (JVP) kernels.
//
// The value part of every operation is the same IEEE operation the FP64
// residual performs, so a tangent kernel reproduces the residual bitwise in
// .v while .d carries the exact directional derivative.  Branches take the
// side selected by the value (the active branch of min/max/|x| and of the
// flux switches).
#pragma once
#include <hip/hip_runtime.h>

namespace fcfd {

struct Dual {
    double v, d;
    __host__ __device__ Dual() : v(0.0), d(0.0) {}
    __host__ __device__ Dual(double a) : v(a), d(0.0) {}
    __host__ __device__ Dual(double a, double b) : v(a), d(b) {}
};

__device__ __forceinline__ double val(const Dual& x) { return x.v; }

__device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator+(const Dual& a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, const Dual& b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(const Dual& a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, const Dual& b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
    double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, const Dual& b) {
    double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ Dual& operator+=(Dual& a, const Dual& b) { a = a + b; return a; }
__device__ __forceinline__ Dual& operator-=(Dual& a, const Dual& b) { a = a - b; return a; }
__device__ __forceinline__ Dual& operator*=(Dual& a, const Dual& b) { a = a * b; return a; }

// At x = 0 the derivative of sqrt is unbounded; the code takes sqrt(0) only
// for a magnitude |V| of a vector that is exactly zero (the no-slip wall face
// of Steger-Warming), where x = V.V has a zero tangent: the tangent 0 is then
// the exact derivative of |V|^2 and the zero subgradient of |V|.
__device__ __forceinline__ Dual sq_root(const Dual& x) {
    double s = sqrt(x.v);
    return Dual(s, s > 0.0 ? x.d / (2.0 * s) : 0.0);
}
__device__ __forceinline__ Dual powr(const Dual& x, double y) {   // y constant
    double p = pow(x.v, y);
    return Dual(p, y * pow(x.v, y - 1.0) * x.d);
}
__device__ __forceinline__ Dual absv(const Dual& x) { return Dual(fabs(x.v), x.v < 0.0 ? -x.d : x.d); }

}  // namespace fcfd
