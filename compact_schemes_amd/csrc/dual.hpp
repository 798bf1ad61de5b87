// dual.hpp -- forward-mode dual numbers for the pointwise Jacobian-transpose
// products of the adjoint: a value and its derivatives along ND seeded
// directions.  Only the operations the pointwise physics uses.
#pragma once

#include <hip/hip_runtime.h>

namespace cs {

using ::sqrt;   // the dual overload below must not hide the scalar ones

template <int ND>
struct Dual {
    double v;
    double d[ND];
    __host__ __device__ Dual() : v(0.0) {
        for (int k = 0; k < ND; k++) d[k] = 0.0;
    }
    __host__ __device__ Dual(double x) : v(x) {       // a constant
        for (int k = 0; k < ND; k++) d[k] = 0.0;
    }
    __host__ __device__ static Dual seed(double x, int dir) {
        Dual r(x);
        r.d[dir] = 1.0;
        return r;
    }
};

template <int ND>
__host__ __device__ inline Dual<ND> operator+(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r;
    r.v = a.v + b.v;
    for (int k = 0; k < ND; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int ND>
__host__ __device__ inline Dual<ND> operator-(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r;
    r.v = a.v - b.v;
    for (int k = 0; k < ND; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int ND>
__host__ __device__ inline Dual<ND> operator-(const Dual<ND>& a) {
    Dual<ND> r;
    r.v = -a.v;
    for (int k = 0; k < ND; k++) r.d[k] = -a.d[k];
    return r;
}
template <int ND>
__host__ __device__ inline Dual<ND> operator*(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r;
    r.v = a.v * b.v;
    for (int k = 0; k < ND; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int ND>
__host__ __device__ inline Dual<ND> operator/(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r;
    const double ib = 1.0 / b.v;
    r.v = a.v * ib;
    for (int k = 0; k < ND; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int ND> __host__ __device__ inline Dual<ND> operator+(const Dual<ND>& a, double b) { return a + Dual<ND>(b); }
template <int ND> __host__ __device__ inline Dual<ND> operator+(double a, const Dual<ND>& b) { return Dual<ND>(a) + b; }
template <int ND> __host__ __device__ inline Dual<ND> operator-(const Dual<ND>& a, double b) { return a - Dual<ND>(b); }
template <int ND> __host__ __device__ inline Dual<ND> operator-(double a, const Dual<ND>& b) { return Dual<ND>(a) - b; }
template <int ND> __host__ __device__ inline Dual<ND> operator*(const Dual<ND>& a, double b) {
    Dual<ND> r;
    r.v = a.v * b;
    for (int k = 0; k < ND; k++) r.d[k] = a.d[k] * b;
    return r;
}
template <int ND> __host__ __device__ inline Dual<ND> operator*(double a, const Dual<ND>& b) { return b * a; }
template <int ND> __host__ __device__ inline Dual<ND> operator/(const Dual<ND>& a, double b) { return a * (1.0 / b); }
template <int ND> __host__ __device__ inline Dual<ND> operator/(double a, const Dual<ND>& b) { return Dual<ND>(a) / b; }
template <int ND> __host__ __device__ inline Dual<ND>& operator+=(Dual<ND>& a, const Dual<ND>& b) { a = a + b; return a; }
template <int ND> __host__ __device__ inline Dual<ND>& operator-=(Dual<ND>& a, const Dual<ND>& b) { a = a - b; return a; }
template <int ND>
__host__ __device__ inline Dual<ND> sqrt(const Dual<ND>& a) {
    Dual<ND> r;
    r.v = ::sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int k = 0; k < ND; k++) r.d[k] = a.d[k] * h;
    return r;
}

__host__ __device__ inline double value_of(double x) { return x; }
template <int ND>
__host__ __device__ inline double value_of(const Dual<ND>& x) { return x.v; }

}  // namespace cs
