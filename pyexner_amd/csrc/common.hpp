// Shared device helpers for the PyExner gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pyexner_hip.h"

namespace pxh {

// utils/constants.py:26-33
constexpr double DRY_TOL = 1e-4;
constexpr double SED_TOL = 1e-5;
constexpr double VEL_TOL = 1e-6;
constexpr double GRAV = 9.81;
// DRY_TOL*0.01 and SED_TOL*0.01 as Python evaluates them (roe_exner.py:421,620-624)
constexpr double DRY_FLOOR = 1.0000000000000002e-06;
constexpr double SED_FLOOR = 1.0000000000000001e-07;

// NaN-propagating max/min, as jnp.maximum / jnp.minimum.
template <typename T>
__device__ __forceinline__ T pmax(T a, T b) {
    return (a != a) ? a : ((b != b) ? b : (a > b ? a : b));
}
template <typename T>
__device__ __forceinline__ T pmin(T a, T b) {
    return (a != a) ? a : ((b != b) ? b : (a < b ? a : b));
}

template <typename T>
__device__ __forceinline__ T sq(T x) { return x * x; }          // lax.integer_pow(x, 2)
template <typename T>
__device__ __forceinline__ T cube(T x) { return x * (x * x); }  // lax.integer_pow(x, 3)

__device__ __forceinline__ float d_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float d_pow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double d_pow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float d_cbrt(float x) { return cbrtf(x); }
__device__ __forceinline__ double d_cbrt(double x) { return cbrt(x); }
__device__ __forceinline__ float d_abs(float x) { return fabsf(x); }
__device__ __forceinline__ double d_abs(double x) { return fabs(x); }

template <typename T>
__device__ __forceinline__ T inf_v() { return __builtin_huge_val(); }
template <>
__device__ __forceinline__ float inf_v<float>() { return __builtin_huge_valf(); }

// Cast a Python-evaluated double constant to the working dtype, as JAX's weak
// typing does when it meets a float32 array.
template <typename T>
__device__ __forceinline__ T C(double v) { return static_cast<T>(v); }

// Typed view of pxh_fields.
template <typename T>
struct F {
    T *h, *hu, *hv, *z, *z_b, *n, *n_b, *G;
    T *z_b_lo;  // compensated-bed low part (float32 only) or null
};

// Global-memory access with a 32-bit element offset.  Field pointers reach the
// kernels inside structs, where the compiler cannot see they are global, and
// would otherwise emit flat loads (which share the LDS wait counter with the
// cross-lane shuffles) and 64-bit address arithmetic.
#define PXH_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ld(const T *p, unsigned i) {
    return ((const PXH_GLOBAL T *)p)[i];
}
template <typename T>
__device__ __forceinline__ void st(T *p, unsigned i, T v) {
    ((PXH_GLOBAL T *)p)[i] = v;
}
// Row-based access for sweeps where every lane of a wavefront is on the same
// row: the row offset is wavefront-uniform (scalar registers, any size), the
// column a 32-bit per-lane offset.
__device__ __forceinline__ unsigned long uni64(unsigned long a) {
    const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(a));
    const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(a >> 32));
    return (static_cast<unsigned long>(hi) << 32) | lo;
}
// row_off must be the same in every active lane of the wavefront.
template <typename T>
__device__ __forceinline__ T ldr(const T *p, long row_off, unsigned col) {
    const unsigned long base = uni64(reinterpret_cast<unsigned long>(p + row_off));
    return *(const PXH_GLOBAL T *)((const PXH_GLOBAL char *)base + col * unsigned(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void str(T *p, long row_off, unsigned col, T v) {
    const unsigned long base = uni64(reinterpret_cast<unsigned long>(p + row_off));
    *(PXH_GLOBAL T *)((PXH_GLOBAL char *)base + col * unsigned(sizeof(T))) = v;
}

inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PXH_OK : static_cast<int>(e);
}

}  // namespace pxh
