// dual.h -- forward-mode derivatives: a value and its derivatives with
// respect to N inputs.  The value of every operation is computed by the same
// double operation as plain code would use, so code templated on the scalar
// type takes the same branches and produces the same values with Dual as
// with double; the derivatives follow by the chain rule.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#pragma clang fp contract(off)

namespace stz {

// the double functions stay visible beside the Dual overloads below
using ::cosh;
using ::exp;
using ::fabs;
using ::sinh;
using ::sqrt;
using ::tanh;

template <int N> struct Dual {
    double v;
    double d[N];

    __host__ __device__ Dual() : v(0) {
        for (int k = 0; k < N; k++) d[k] = 0;
    }
    __host__ __device__ Dual(double x) : v(x) {
        for (int k = 0; k < N; k++) d[k] = 0;
    }
    // input k of the N with value x
    __host__ __device__ static Dual var(double x, int k) {
        Dual r(x);
        r.d[k] = 1;
        return r;
    }
};

template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N> &a) {
    Dual<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; k++) r.d[k] = -a.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1 / b.v;
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N> &a, double b) { return a + Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator+(double a, const Dual<N> &b) { return Dual<N>(a) + b; }
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N> &a, double b) { return a - Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator-(double a, const Dual<N> &b) { return Dual<N>(a) - b; }
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N> &a, double b) {
    Dual<N> r;
    r.v = a.v * b;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator*(double a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a * b.v;
    for (int k = 0; k < N; k++) r.d[k] = a * b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N> &a, double b) { return a / Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator/(double a, const Dual<N> &b) { return Dual<N>(a) / b; }
template <int N> __host__ __device__ inline Dual<N> &operator+=(Dual<N> &a, const Dual<N> &b) { return a = a + b; }
template <int N> __host__ __device__ inline Dual<N> &operator-=(Dual<N> &a, const Dual<N> &b) { return a = a - b; }
template <int N> __host__ __device__ inline Dual<N> &operator/=(Dual<N> &a, const Dual<N> &b) { return a = a / b; }

// comparisons look at values only
template <int N> __host__ __device__ inline bool operator<(const Dual<N> &a, double b) { return a.v < b; }
template <int N> __host__ __device__ inline bool operator>(const Dual<N> &a, double b) { return a.v > b; }
template <int N> __host__ __device__ inline bool operator<(const Dual<N> &a, const Dual<N> &b) { return a.v < b.v; }
template <int N> __host__ __device__ inline bool operator>(const Dual<N> &a, const Dual<N> &b) { return a.v > b.v; }
template <int N> __host__ __device__ inline bool operator>=(double a, const Dual<N> &b) { return a >= b.v; }

template <int N> __host__ __device__ inline Dual<N> chain(const Dual<N> &a, double fv, double dfdx) {
    Dual<N> r;
    r.v = fv;
    for (int k = 0; k < N; k++) r.d[k] = dfdx * a.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> exp(const Dual<N> &a) {
    const double e = ::exp(a.v);
    return chain(a, e, e);
}
template <int N> __host__ __device__ inline Dual<N> sqrt(const Dual<N> &a) {
    const double q = ::sqrt(a.v);
    return chain(a, q, 0.5 / q);
}
template <int N> __host__ __device__ inline Dual<N> cosh(const Dual<N> &a) { return chain(a, ::cosh(a.v), ::sinh(a.v)); }
template <int N> __host__ __device__ inline Dual<N> sinh(const Dual<N> &a) { return chain(a, ::sinh(a.v), ::cosh(a.v)); }
template <int N> __host__ __device__ inline Dual<N> tanh(const Dual<N> &a) {
    const double t = ::tanh(a.v);
    return chain(a, t, 1 - t * t);
}
template <int N> __host__ __device__ inline Dual<N> fabs(const Dual<N> &a) { return a.v < 0 ? -a : a; }

}  // namespace stz
