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

// The sediment table as the kernels take it: the C table plus its float32
// fold for the hardware-math path, fw = 8 frac C/D / (1 - porosity), so that
// G = sum_k fw_k h^(-1/2) sqrt(dtheta_k) dtheta_k / (theta_k^1.5 + 1e-5).
struct SedK : pxh_sediments {
    float fAB[PXH_MAX_SED], fthc[PXH_MAX_SED], fw[PXH_MAX_SED];
};

static inline SedK sed_k(const pxh_sediments &s) {
    SedK k;
    static_cast<pxh_sediments &>(k) = s;
    for (int i = 0; i < PXH_MAX_SED; ++i) {
        const bool on = i < s.n_sed;
        k.fAB[i] = on ? float(s.AB[i]) : 0.0f;
        k.fthc[i] = on ? float(s.theta_c[i]) : 0.0f;
        k.fw[i] = on ? float(8.0 * s.frac[i] * s.CD[i] / s.one_minus_por) : 0.0f;
    }
    return k;
}

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
// The same with a 32-bit row offset (elements) from a base the caller has moved
// to the first row it touches, so that the whole byte offset fits 32 bits: the
// access then takes the base in scalar registers and one vector register for
// the offset, with no 64-bit address arithmetic per load.  The caller keeps
// (row_off + col) * sizeof(T) below 2^32.
template <typename T>
__device__ __forceinline__ T ldr(const T *p, unsigned row_off, unsigned col) {
    const unsigned long base = uni64(reinterpret_cast<unsigned long>(p));
    const unsigned ro = __builtin_amdgcn_readfirstlane(row_off * unsigned(sizeof(T)));
    return *(const PXH_GLOBAL T *)((const PXH_GLOBAL char *)base + (ro + col * unsigned(sizeof(T))));
}
template <typename T>
__device__ __forceinline__ void str(T *p, unsigned row_off, unsigned col, T v) {
    const unsigned long base = uni64(reinterpret_cast<unsigned long>(p));
    const unsigned ro = __builtin_amdgcn_readfirstlane(row_off * unsigned(sizeof(T)));
    *(PXH_GLOBAL T *)((PXH_GLOBAL char *)base + (ro + col * unsigned(sizeof(T)))) = v;
}

// A ring-fill entry (pxh_bc_entry, 16 bytes) read as four words: its fields,
// and the bytes of its handler chain by shifts (an entry struct indexed by a
// loop variable would be copied to scratch memory for every lane).
static_assert(sizeof(pxh_bc_entry) == 16, "a ring-fill entry is read as four words");
struct BcEnt {
    uint4 w;
    __device__ int target() const { return int(w.x); }
    __device__ int src() const { return int(w.y); }
    __device__ unsigned group() const { return w.z & 0xffu; }
    __device__ unsigned nchain() const { return (w.z >> 8) & 0xffu; }
    __device__ unsigned chain(int q) const {  // byte 10 + q of the entry
        const unsigned b = 2u + unsigned(q);
        return ((b < 4u ? w.z : w.w) >> (8u * (b & 3u))) & 0xffu;
    }
    // the entry as the struct, for code that reads its first chain byte only
    __device__ pxh_bc_entry first() const {
        pxh_bc_entry e{};
        e.target = target();
        e.src = src();
        e.group = uint8_t(group());
        e.nchain = uint8_t(nchain());
        e.chain[0] = uint8_t(chain(0));
        return e;
    }
};
__device__ __forceinline__ BcEnt bc_ent(const pxh_bc_entry *ent, long k) {
    return BcEnt{reinterpret_cast<const uint4 *>(ent)[k]};
}

// Four consecutive floats per cell (the packed records of PXH_PACKED_STAGE)
// with the row addressing of ldr / str above: element (row_off + col) of an
// array of four-float records, 32-bit byte offsets from a moved base.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4v ldr4(const float *p, unsigned row_off, unsigned col) {
    const unsigned long base = uni64(reinterpret_cast<unsigned long>(p));
    const unsigned ro = __builtin_amdgcn_readfirstlane(row_off * 16u);
    return *(const PXH_GLOBAL f4v *)((const PXH_GLOBAL char *)base + (ro + col * 16u));
}
__device__ __forceinline__ void str4(float *p, unsigned row_off, unsigned col, f4v v) {
    const unsigned long base = uni64(reinterpret_cast<unsigned long>(p));
    const unsigned ro = __builtin_amdgcn_readfirstlane(row_off * 16u);
    *(PXH_GLOBAL f4v *)((PXH_GLOBAL char *)base + (ro + col * 16u)) = v;
}

// f with every field pointer moved by off elements (null stays null).
template <typename T>
__host__ __device__ __forceinline__ F<T> shifted(const F<T> &f, long off) {
    auto mv = [off](T *p) { return p ? p + off : p; };
    return F<T>{mv(f.h), mv(f.hu), mv(f.hv), mv(f.z), mv(f.z_b), mv(f.n), mv(f.n_b), mv(f.G), mv(f.z_b_lo)};
}

inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PXH_OK : static_cast<int>(e);
}

}  // namespace pxh
