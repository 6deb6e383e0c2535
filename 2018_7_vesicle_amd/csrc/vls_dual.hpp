// Forward-mode dual numbers for the discrete adjoint: a value and N partial
// derivatives.  Every operation computes its value part with the same IEEE
// operation as the double code (a*b, a+b, a/b, fma, sqrt), so a formula
// evaluated on duals takes exactly the branches the double evaluation takes
// and its value part equals the double result bit for bit; the derivative
// parts are the exact derivatives of that formula with the branches frozen.
#pragma once
#include <hip/hip_runtime.h>

namespace vls {
// own namespace: the double code of namespace vls keeps finding ::fma and
// ::sqrt; calls on duals find these overloads by argument-dependent lookup
namespace ad {

template <int N>
struct Dual {
    double v;
    double d[N];
    __device__ Dual() {}
    __device__ explicit Dual(double x) : v(x)
    {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(double v)
{
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}

// the k-th input variable with value v
template <int N>
__device__ __forceinline__ Dual<N> dvar(double v, int k)
{
    Dual<N> r = dconst<N>(v);
    r.d[k] = 1.0;
    return r;
}

template <int N>
__device__ __forceinline__ double val(const Dual<N>& a)
{
    return a.v;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a)
{
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
// mixed forms with a constant
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b)
{
    Dual<N> r = b;
    r.v = a + b.v;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v / b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b)
{
    return dconst<N>(a) / b;
}

template <int N>
__device__ __forceinline__ Dual<N> fma(const Dual<N>& a, const Dual<N>& b, const Dual<N>& c)
{
    Dual<N> r;
    r.v = ::fma(a.v, b.v, c.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k] + c.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> fma(double a, const Dual<N>& b, const Dual<N>& c)
{
    Dual<N> r;
    r.v = ::fma(a, b.v, c.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k] + c.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> sqrt(const Dual<N>& a)
{
    Dual<N> r;
    r.v = ::sqrt(a.v);
    const double h = (r.v > 0) ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> dmax(const Dual<N>& a, const Dual<N>& b)
{
    return (a.v < b.v) ? b : a;
}
template <int N>
__device__ __forceinline__ Dual<N> dmin(const Dual<N>& a, const Dual<N>& b)
{
    return (a.v < b.v) ? a : b;
}
template <int N>
__device__ __forceinline__ Dual<N> sq(const Dual<N>& a)
{
    return a * a;
}

}  // namespace ad
using ad::Dual;
}  // namespace vls
