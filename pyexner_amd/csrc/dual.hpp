// Forward-mode dual numbers (value and N first derivatives) in double, for the
// local Jacobians of the adjoint kernels: the edge functions of
// flux_roe_exner.hpp are templates in their arithmetic type, so evaluating one
// with Dual<N> arguments gives its derivatives along N seeded directions with
// the same branches the forward took.
//
// Branch tests compare values only.  sqrt, cbrt and pow at 0 take derivative
// 0 (the subgradient where the forward clamps to 0 before the root, e.g. the
// excess Shields stress below critical), which keeps an untaken or clamped
// branch from turning 0 * inf into NaN.
#pragma once
#include "common.hpp"

namespace pxh {

template <int N>
struct Dual {
    double v;
    double d[N];
    __host__ __device__ Dual() : v(0.0) {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    __host__ __device__ Dual(double x) : v(x) {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    __host__ __device__ explicit operator double() const { return v; }
    // seeded variable: value x, derivative 1 along direction k
    __host__ __device__ static Dual var(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
    __host__ __device__ Dual &operator+=(const Dual &o) {
        v += o.v;
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] += o.d[k];
        return *this;
    }
};

template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
#define PXH_DUAL_MIXED(OP)                                                                            \
    template <int N>                                                                                  \
    __host__ __device__ __forceinline__ Dual<N> operator OP(const Dual<N> &a, double b) {             \
        return a OP Dual<N>(b);                                                                       \
    }                                                                                                 \
    template <int N>                                                                                  \
    __host__ __device__ __forceinline__ Dual<N> operator OP(double a, const Dual<N> &b) {             \
        return Dual<N>(a) OP b;                                                                       \
    }
PXH_DUAL_MIXED(+)
PXH_DUAL_MIXED(-)
PXH_DUAL_MIXED(*)
PXH_DUAL_MIXED(/)
#undef PXH_DUAL_MIXED

#define PXH_DUAL_CMP(OP)                                                                                          \
    template <int N>                                                                                              \
    __host__ __device__ __forceinline__ bool operator OP(const Dual<N> &a, const Dual<N> &b) { return a.v OP b.v; } \
    template <int N>                                                                                              \
    __host__ __device__ __forceinline__ bool operator OP(const Dual<N> &a, double b) { return a.v OP b; }          \
    template <int N>                                                                                              \
    __host__ __device__ __forceinline__ bool operator OP(double a, const Dual<N> &b) { return a OP b.v; }
PXH_DUAL_CMP(<)
PXH_DUAL_CMP(>)
PXH_DUAL_CMP(<=)
PXH_DUAL_CMP(>=)
PXH_DUAL_CMP(==)
PXH_DUAL_CMP(!=)
#undef PXH_DUAL_CMP

// f(a) with f'(a) = fp (the chain rule for a unary function).
template <int N>
__host__ __device__ __forceinline__ Dual<N> dual_unary(const Dual<N> &a, double f, double fp) {
    Dual<N> r;
    r.v = f;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = fp * a.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> d_sqrt(const Dual<N> &a) {
    const double s = __builtin_sqrt(a.v);
    return dual_unary(a, s, s > 0.0 ? 0.5 / s : 0.0);
}
template <int N>
__device__ __forceinline__ Dual<N> d_cbrt(const Dual<N> &a) {
    const double c = cbrt(a.v);
    return dual_unary(a, c, c != 0.0 ? 1.0 / (3.0 * c * c) : 0.0);
}
// power with a constant exponent (the forward only raises to 4/3 and 1.5)
template <int N>
__device__ __forceinline__ Dual<N> d_pow(const Dual<N> &a, const Dual<N> &e) {
    const double p = pow(a.v, e.v);
    return dual_unary(a, p, a.v != 0.0 ? e.v * pow(a.v, e.v - 1.0) : 0.0);
}
template <int N>
__device__ __forceinline__ Dual<N> d_abs(const Dual<N> &a) {
    return dual_unary(a, __builtin_fabs(a.v), a.v < 0.0 ? -1.0 : 1.0);
}

}  // namespace pxh
