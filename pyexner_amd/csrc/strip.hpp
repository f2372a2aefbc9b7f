// Wavefront-strip sweeps for the Roe-Exner stage.
//
// A 64-lane wavefront owns a strip of 62 output columns (lanes 1..62; lanes 0
// and 63 are the west/east halo columns) and walks down a strip of rows
// keeping the current and next row in registers, with the two rows after
// them in flight.  Every edge flux is evaluated once: the x edge (c | c+1) by
// lane c with lane c+1's values from a cross-lane DPP shift (its upwP reaches
// lane c+1 by the reverse shift), the y edge (r | r+1) by the lane that owns
// the column, carried to the next row in a register.  Row loads are
// coalesced and addressed from a wavefront-uniform row base; there is no
// integer division.
//
// FAST selects the float32 arithmetic: hardware reciprocal, reciprocal square
// root, square root, log2 and exp2 (about 1 ulp each) with reciprocals shared
// across an edge, and the CFL cubic in float32.  FAST = false keeps the
// reference's operation sequence with IEEE division and square root (the
// double build, and float32 when exact agreement with a float32 restatement is
// wanted).
#pragma once
#include "common.hpp"
#include "flux_roe_exner.hpp"

namespace pxh {

constexpr int SW = 64;         // lanes per strip
constexpr int SO_WIDE = 62;    // output columns per strip (lanes 1..62)
constexpr int SR_HYDRO = 16;   // rows per strip, hydro sweep
constexpr int SR_BED = 16;     // rows per strip, bed and dt sweeps
constexpr int SWPB = 2;        // strips (wavefronts) per block, stacked in y

// Strip geometry: lanes 1..SO own columns SO s .. SO s + SO - 1; lane 0 holds
// the west neighbour column and lane SO + 1 the east one.  With SO = 62 the
// row stores of two neighbouring strips split the 64-byte block at their
// common edge; the XCD-aware block order below keeps the two wavefronts on
// one L2, where the halves merge (split writes that reach HBM separately cost
// about 40% of a store-heavy sweep).  The strip heights and two wavefronts
// per block were chosen by timing the SSPRK2 step on an 8192 x 8192 grid.
struct StripPos {
    int lane, c, cc, r0, r1;
    bool out;
};

// XCD-aware block order.  The hardware hands linear block L to XCD L % 8;
// XCD k takes the logical blocks [k n / 8, (k + 1) n / 8) in order, a band of
// strip rows, so the strips on either side of a shared column edge run on one
// XCD (one L2) at nearly the same time, and their partial writes of the
// 64-byte blocks at the edge merge in that L2 instead of reaching HBM apart.
constexpr int NXCD = 8;
__device__ __forceinline__ void xcd_block(int &bx, int &by) {
    const int gx = gridDim.x, n = gx * gridDim.y;
    const int L = blockIdx.x + gx * blockIdx.y;
    const int k = L % NXCD, j = L / NXCD, q = n / NXCD, rm = n % NXCD;
    const int lb = k * q + (k < rm ? k : rm) + j;
    bx = lb % gx;
    by = lb / gx;
}

// The seam-column buffer of a combine sweep: one record of 10 values per strip
// column bx and row r (records of a strip column in row order): the combined
// h, hu, hv, G of the strip's first owned column (side 0) at [0, 4), of its
// last (side 1) at [4, 8), and their nodata flags (0 or 1) at 8 and 9.  The
// edges between two strips' columns are reduced from it after the sweep.  A
// row's values sit together, so the sweep's stores of one row fill one span
// instead of ten scattered cache lines.
constexpr int SEAM_REC = 10;
template <typename T>
__host__ __device__ __forceinline__ T *seam_rec(T *xs, int bx, int r, int ny) {
    return xs + (long(bx) * ny + r) * SEAM_REC;
}
// One side of seam record q.
template <typename T>
__device__ __forceinline__ void seam_put(T *q, int side, T h, T hu, T hv, T G, bool nod) {
    q[4 * side] = h;
    q[4 * side + 1] = hu;
    q[4 * side + 2] = hv;
    q[4 * side + 3] = G;
    q[8 + side] = nod ? T(1) : T(0);
}

template <int SO, int SROWS>
__device__ __forceinline__ StripPos strip_pos_at(int bx, int by, int ny, int nx) {
    StripPos p;
    p.lane = threadIdx.x & (SW - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / SW);  // uniform: rows are scalar
    p.c = bx * SO - 1 + p.lane;
    p.cc = p.c < 0 ? 0 : (p.c >= nx ? nx - 1 : p.c);
    p.out = p.lane >= 1 && p.lane <= SO && p.c < nx;
    p.r0 = (by * SWPB + w) * SROWS;
    p.r1 = p.r0 + SROWS < ny ? p.r0 + SROWS : ny;
    return p;
}

template <int SO, int SROWS>
__device__ __forceinline__ StripPos strip_pos(int ny, int nx) {
    int bx, by;
    xcd_block(bx, by);
    return strip_pos_at<SO, SROWS>(bx, by, ny, nx);
}

template <int SO, int SROWS>
static inline dim3 strip_grid(int ny, int nx) {
    const int sx = (nx + SO - 1) / SO;
    const int sy = (ny + SROWS - 1) / SROWS;
    return dim3(sx, (sy + SWPB - 1) / SWPB);
}

// Wait for every outstanding vector-memory load (vmcnt(0); expcnt and lgkmcnt
// left free).  Issued once before a row loop whose software pipeline keeps
// loads in flight across iterations: with the prologue's loads drained, the
// waits inside the loop only cover the loads the loop itself issued.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Whole-wavefront shifts by one lane on the DPP path (no LDS round trip):
// wave_shl:1 gives lane i the value of lane i+1, wave_shr:1 that of lane i-1;
// the lane with no source (63, resp. 0) reads 0 (bound_ctrl).  Those are the
// strips' halo lanes, whose shifted-in values no output uses, and the zero
// fill lets the compiler fold the shift into the instruction that consumes it.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true); }
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_shift(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lane values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp_i32<CTRL>(__builtin_bit_cast(int, v)));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = dpp_i32<CTRL>(int(b & 0xffffffffLL));
        const int hi = dpp_i32<CTRL>(int(b >> 32));
        return __builtin_bit_cast(T, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    }
}
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
template <typename T>
__device__ __forceinline__ T shdn(T v) { return dpp_shift<DPP_WAVE_SHL1>(v); }
template <typename T>
__device__ __forceinline__ T shup(T v) { return dpp_shift<DPP_WAVE_SHR1>(v); }

// ---------------------------------------------------------------------------
// arithmetic policies
// ---------------------------------------------------------------------------
template <typename T, bool FAST>
struct Mth {  // exact: the reference's operations
    static __device__ __forceinline__ T rcp(T x) { return T(1) / x; }
    static __device__ __forceinline__ T div(T a, T b) { return a / b; }
    static __device__ __forceinline__ T sqrt(T x) { return d_sqrt(x); }
};
template <>
struct Mth<float, true> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float log2(float x) { return __builtin_amdgcn_logf(x); }
    static __device__ __forceinline__ float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
};

// ---------------------------------------------------------------------------
// hydro + update + momentum corrections
// ---------------------------------------------------------------------------
template <typename T>
struct HD {          // derived per-cell values for the Roe-Exner edge
    T h, hu, hv;     // raw state
    T hf, u, v;      // floored depth, velocities (0 when dry)
    T s, e;          // sqrt(hf), sqrt(g hf)
    T n;             // n_b
    T zt;            // z + floored z_b (bed thrust)
    T zr;            // z + z_b (momentum corrections)
    T zb;            // raw z_b (the fused stage's bed update)
    int flags;       // bit0 wet (h > DRY_TOL), bit1 nodata
};

template <typename T>
struct HR {          // raw per-cell inputs of the hydro sweep
    T h, hu, hv, zb, z, n;
    int m;
};

// RO: long (absolute row offset) or unsigned (from a moved base, 32-bit).
template <typename T, typename RO>
__device__ __forceinline__ HR<T> hr_load(const F<T> &s, const uint8_t *m0, RO ro, unsigned c) {
    HR<T> q;
    q.h = ldr(s.h, ro, c); q.hu = ldr(s.hu, ro, c); q.hv = ldr(s.hv, ro, c);
    q.zb = ldr(s.z_b, ro, c); q.z = ldr(s.z, ro, c); q.n = ldr(s.n_b, ro, c);
    q.m = ldr(m0, ro, c);
    return q;
}

// hr_load on a grid with no nodata cell (PXH_NO_NODATA): the mask is zero
// everywhere, so it is not loaded.
template <typename T, typename RO>
__device__ __forceinline__ HR<T> hr_load_no_nodata(const F<T> &s, RO ro, unsigned c) {
    HR<T> q;
    q.h = ldr(s.h, ro, c); q.hu = ldr(s.hu, ro, c); q.hv = ldr(s.hv, ro, c);
    q.zb = ldr(s.z_b, ro, c); q.z = ldr(s.z, ro, c); q.n = ldr(s.n_b, ro, c);
    q.m = 0;
    return q;
}

// hr_load on a grid with no nodata cell whose fixed bed z holds one value
// zu in every cell (PXH_NO_NODATA with PXH_Z_UNIFORM): neither the mask nor z
// is loaded.
template <typename T, typename RO>
__device__ __forceinline__ HR<T> hr_load_static(const F<T> &s, RO ro, unsigned c, T zu) {
    HR<T> q;
    q.h = ldr(s.h, ro, c); q.hu = ldr(s.hu, ro, c); q.hv = ldr(s.hv, ro, c);
    q.zb = ldr(s.z_b, ro, c); q.z = zu; q.n = ldr(s.n_b, ro, c);
    q.m = 0;
    return q;
}

template <typename T, bool FAST>
__device__ __forceinline__ HD<T> hd_derive(const HR<T> &q) {
    const T T_ = C<T>(DRY_TOL), S_ = C<T>(SED_TOL), g = C<T>(GRAV), zero = T(0);
    HD<T> d;
    d.h = q.h; d.hu = q.hu; d.hv = q.hv;
    d.n = q.n;
    d.hf = pmax(d.h, T_);
    const bool wet = d.hf > T_;
    if constexpr (FAST) {
        using M = Mth<float, true>;
        const float r = M::rsq(d.hf);  // 1 / sqrt(hf): gives 1 / hf and sqrt(hf)
        const float ih = r * r;
        d.u = wet ? d.hu * ih : zero;
        d.v = wet ? d.hv * ih : zero;
        d.s = d.hf * r;
        d.e = d.s * 3.1320919526731650f;  // sqrt(9.81)
    } else {
        d.u = wet ? d.hu / d.hf : zero;
        d.v = wet ? d.hv / d.hf : zero;
        d.s = d_sqrt(d.hf);
        d.e = d_sqrt(g * d.hf);
    }
    d.zt = q.z + (q.zb > S_ ? q.zb : C<T>(SED_FLOOR));
    d.zr = q.z + q.zb;
    d.zb = q.zb;
    d.flags = (d.h > T_ ? 1 : 0) | (q.m ? 2 : 0);
    return d;
}

template <typename T, bool FAST, typename RO>
__device__ __forceinline__ HD<T> hd_load(const F<T> &s, const uint8_t *m0, RO ro, unsigned c) {
    return hd_derive<T, FAST>(hr_load(s, m0, ro, c));
}

// The same for a cell whose row may differ from lane to lane (per-lane 64-bit
// element index; no wavefront-uniform row base).
template <typename T, bool FAST>
__device__ __forceinline__ HD<T> hd_load_lane(const F<T> &s, const uint8_t *m0, long i) {
    HR<T> q;
    q.h = ((const PXH_GLOBAL T *)s.h)[i]; q.hu = ((const PXH_GLOBAL T *)s.hu)[i];
    q.hv = ((const PXH_GLOBAL T *)s.hv)[i]; q.zb = ((const PXH_GLOBAL T *)s.z_b)[i];
    q.z = ((const PXH_GLOBAL T *)s.z)[i]; q.n = ((const PXH_GLOBAL T *)s.n_b)[i];
    q.m = ((const PXH_GLOBAL uint8_t *)m0)[i];
    return hd_derive<T, FAST>(q);
}

// The packed form of a first SSPRK2 stage's output (PXH_PACKED_STAGE, float
// only): h, hu, hv, z_b of cell i at q[4 i .. 4 i + 3]; the static fields and
// the mask as in s.  NN: no nodata cell (the mask is not loaded); ZU: z holds
// the one value zu (not loaded).
template <bool NN, bool ZU, typename T>
__device__ __forceinline__ HR<T> hr_load_q(const F<T> &s, const T *q, const uint8_t *m0, unsigned ro, unsigned c,
                                           T zu) {
    static_assert(sizeof(T) == 4, "packed records hold float32");
    HR<T> r;
    const f4v v = ldr4(reinterpret_cast<const float *>(q), ro, c);
    r.h = v.x; r.hu = v.y; r.hv = v.z; r.zb = v.w;
    r.z = ZU ? zu : ldr(s.z, ro, c);
    r.n = ldr(s.n_b, ro, c);
    r.m = NN ? 0 : ldr(m0, ro, c);
    return r;
}

template <typename T, bool FAST>
__device__ __forceinline__ HD<T> hd_load_lane_q(const F<T> &s, const T *q, const uint8_t *m0, long i) {
    static_assert(sizeof(T) == 4, "packed records hold float32");
    HR<T> r;
    const f4v v = *(const PXH_GLOBAL f4v *)(reinterpret_cast<const float *>(q) + 4 * i);
    r.h = v.x; r.hu = v.y; r.hv = v.z; r.zb = v.w;
    r.z = ((const PXH_GLOBAL T *)s.z)[i]; r.n = ((const PXH_GLOBAL T *)s.n_b)[i];
    r.m = ((const PXH_GLOBAL uint8_t *)m0)[i];
    return hd_derive<T, FAST>(r);
}

template <typename T>
__device__ __forceinline__ HD<T> hd_shdn(const HD<T> &a) {
    HD<T> b;
    b.h = shdn(a.h); b.hu = shdn(a.hu); b.hv = shdn(a.hv);
    b.hf = shdn(a.hf); b.u = shdn(a.u); b.v = shdn(a.v);
    b.s = shdn(a.s); b.e = shdn(a.e); b.n = shdn(a.n);
    b.zt = shdn(a.zt); b.zr = shdn(a.zr); b.flags = shdn(a.flags);
    return b;
}

template <typename T>
__device__ __forceinline__ bool hd_active(const HD<T> &a, const HD<T> &b) {
    return ((a.flags & 1) || (b.flags & 1)) && !((a.flags | b.flags) & 2);
}

// The Roe-Exner upwinding (roe_exner.py:484-531) of the three waves and their
// entropy parts; P to cell j, M to cell i, rotated back to the global frame.
template <int AX, typename T>
__device__ __forceinline__ void rx_upwind(T l1, T l2, T l3, T lE1, T lE3, T a1, T a2, T a3, T b1, T vt, T ct,
                                          T P[3], T M[3]) {
    const T zero = T(0);
    T p0 = zero, p1 = zero, p2 = zero, m0 = zero, m1 = zero, m2 = zero;
    auto acc = [&](T lam, T f0, T f1, T f2) {
        if (lam > zero) { p0 += f0; p1 += f1; p2 += f2; }
        if (lam <= zero) { m0 += f0; m1 += f1; m2 += f2; }
    };
    { const T f = l1 * a1 - b1; acc(l1, f * T(1), f * l1, f * vt); }
    { const T f = l2 * a2 - zero; acc(l2, f * zero, f * zero, f * ct); }
    { const T f = l3 * a3 - (-b1); acc(l3, f * T(1), f * l3, f * vt); }
    { const T f = lE1 * a1; acc(lE1, f * T(1), f * l1, f * vt); }
    { const T f = zero * a2; acc(zero, f * zero, f * zero, f * ct); }
    { const T f = lE3 * a3; acc(lE3, f * T(1), f * l3, f * vt); }
    unrot<AX>(p1, p2);
    unrot<AX>(m1, m2);
    P[0] = p0; P[1] = p1; P[2] = p2;
    M[0] = m0; M[1] = m1; M[2] = m2;
}

// roe_solver (roe_exner.py:312-533) on derived values.  Exact form: the same
// operations on the same values as roe_exner_edge.
template <int AX, typename T>
__device__ __forceinline__ void rx_edge_exact(const HD<T> &I, const HD<T> &J, T dx, T P[3], T M[3]) {
    const T T_ = C<T>(DRY_TOL), g = C<T>(GRAV), V = C<T>(VEL_TOL), zero = T(0);
    const T hi = I.hf, hj = J.hf;
    T uhi, vhi, uhj, vhj;
    rot<AX>(I.u, I.v, uhi, vhi);
    rot<AX>(J.u, J.v, uhj, vhj);
    const T si = I.s, sj = J.s;
    const T den = si + sj;
    const T ht = T(0.5) * (hi + hj);
    const T ct = d_sqrt(g * ht);
    const T ut = (uhi * si + uhj * sj) / den;
    const T vt = (vhi * si + vhj * sj) / den;
    const T l1r = ut - ct, l2 = ut, l3r = ut + ct;
    T ei = uhi - I.e, ej = uhj - J.e;
    T lE1 = zero, l1 = l1r;
    if ((ei < zero) && (ej > zero)) {
        lE1 = l1r - ei * (ej - l1r) / (ej - ei);
        l1 = ei * (ej - l1r) / (ej - ei);
    }
    ei = uhi + I.e;
    ej = uhj + J.e;
    T lE3 = zero, l3 = l3r;
    if ((ei < zero) && (ej > zero)) {
        lE3 = l3r - ej * (l3r - ei) / (ej - ei);
        l3 = ej * (l3r - ei) / ((ej - ei) + V);
    }
    const T dh = hj - hi;
    const T dhu = hj * uhj - hi * uhi;
    const T dhv = hj * vhj - hi * vhi;
    const T dn2 = T(-2) * ct;
    const T a1 = (-l3r * dh + dhu) / dn2;
    const T a2 = (-vt * dh + dhv) / ct;
    const T a3 = (l1r * dh - dhu) / dn2;
    const T dz = J.zt - I.zt;
    const T thrust = (-g * ht) * dz;
    const T nt = T(0.5) * (I.n + J.n);
    T tau = zero;
    if (nt != zero) {
        const T sf = (sq(nt) * d_sqrt(sq(ut) + sq(vt)) * ut) / pmax(T_, d_pow(ht, C<T>(4.0 / 3.0)));
        tau = g * ht * sf * dx;
    }
    const T b1 = (thrust - tau) / dn2;
    rx_upwind<AX>(l1, l2, l3, lE1, lE3, a1, a2, a3, b1, vt, ct, P, M);
}

// Upwinding by eigenvector coefficients: every wave's flux lies along one
// column of P (r1 = (1, l1, vt), r2 = (0, 0, ct), r3 = (1, l3, vt)), so the
// coefficients going to each side are summed first and the two vectors are
// assembled once.  Mathematically the same split as rx_upwind, with a
// different summation order (float32 hardware path only).  The products and
// sums are written as explicit fused multiply-adds so that the subcritical
// path below rounds exactly as this one does.
template <int AX>
__device__ __forceinline__ void rx_upwind_coef(float l1, float l2, float l3, float lE1, float lE3, float a1,
                                               float a2, float a3, float b1, float vt, float ct, float P[3],
                                               float M[3]) {
    const float zero = 0.0f;
    const float f1 = __builtin_fmaf(l1, a1, -b1), f2 = l2 * a2, f3 = __builtin_fmaf(l3, a3, b1);
    const float e1 = lE1 * a1, e3 = lE3 * a3;
    // a NaN speed sends its wave to neither side, as in rx_upwind
    const float p1 = (l1 > zero ? f1 : zero) + (lE1 > zero ? e1 : zero);
    const float m1 = (l1 <= zero ? f1 : zero) + (lE1 <= zero ? e1 : zero);
    const float p3 = (l3 > zero ? f3 : zero) + (lE3 > zero ? e3 : zero);
    const float m3 = (l3 <= zero ? f3 : zero) + (lE3 <= zero ? e3 : zero);
    const float p2 = l2 > zero ? f2 : zero;
    const float m2 = l2 <= zero ? f2 : zero;
    float P1 = __builtin_fmaf(p1, l1, p3 * l3), M1 = __builtin_fmaf(m3, l3, m1 * l1);
    const float P0 = p1 + p3, M0 = m1 + m3;
    float P2 = __builtin_fmaf(P0, vt, p2 * ct), M2 = __builtin_fmaf(M0, vt, m2 * ct);
    unrot<AX>(P1, P2);
    unrot<AX>(M1, M2);
    P[0] = P0; P[1] = P1; P[2] = P2;
    M[0] = M0; M[1] = M1; M[2] = M2;
}

// The same edge in float32 hardware arithmetic: one reciprocal for the Roe
// average, one reciprocal square root for the celerity and its inverse, and
// h~^(-4/3) as exp2(-4/3 log2 h~).
//
// Subcritical edges (l1 <= 0 < l3) without an entropy fix (no transonic
// rarefaction in wave 1 or 3) send wave 1 to M and wave 3 to P whatever the
// data, and carry no entropy parts: when every lane whose result is used
// (care) has such an edge, the wavefront takes the short form, which gives
// rx_upwind_coef's values bit for bit (the parts it drops are exact zeros).
// The choice is per wavefront, so it never changes a result.
template <int AX>
__device__ __forceinline__ void rx_edge_fast(const HD<float> &I, const HD<float> &J, float dx, float P[3],
                                             float M[3], bool care = true) {
#pragma clang fp contract(on)
    using Mf = Mth<float, true>;
    const float T_ = float(DRY_TOL), g = float(GRAV), V = float(VEL_TOL), zero = 0.0f;
    const float hi = I.hf, hj = J.hf;
    float uhi, vhi, uhj, vhj;
    rot<AX>(I.u, I.v, uhi, vhi);
    rot<AX>(J.u, J.v, uhj, vhj);
    const float si = I.s, sj = J.s;
    const float iden = Mf::rcp(si + sj);
    const float ht = 0.5f * (hi + hj);
    const float gh = g * ht;
    const float ict = Mf::rsq(gh);
    const float ct = gh * ict;
    const float ut = (uhi * si + uhj * sj) * iden;
    const float vt = (vhi * si + vhj * sj) * iden;
    const float l1r = ut - ct, l2 = ut, l3r = ut + ct;
    const float ei1 = uhi - I.e, ej1 = uhj - J.e;
    const float ei3 = uhi + I.e, ej3 = uhj + J.e;
    const bool hh1 = (ei1 < zero) && (ej1 > zero);
    const bool hh3 = (ei3 < zero) && (ej3 > zero);
    const float dh = hj - hi;
    const float dhu = hj * uhj - hi * uhi;
    const float dhv = hj * vhj - hi * vhi;
    const float ih2 = -0.5f * ict;  // 1 / (-2 c~)
    const float a1 = (-l3r * dh + dhu) * ih2;
    const float a2 = (-vt * dh + dhv) * ict;
    const float a3 = (l1r * dh - dhu) * ih2;
    const float dz = J.zt - I.zt;
    float src = (-g * ht) * dz;
    const float nt = 0.5f * (I.n + J.n);
    if (nt != zero) {
        // 1 / max(DRY_TOL, h~^(4/3)) = min(1 / DRY_TOL, h~^(-4/3))
        const float ih43 = pmin(1.0f / T_, Mf::exp2(-1.3333333333333333f * Mf::log2(ht)));
        const float sf = (nt * nt * Mf::sqrt(ut * ut + vt * vt) * ut) * ih43;
        src = src - g * ht * sf * dx;
    }
    const float b1 = src * ih2;
    const bool plain = !hh1 && !hh3 && !(l1r > zero) && (l3r > zero);
    if (__builtin_amdgcn_ballot_w64(care && !plain) == 0) {
        const float f1 = __builtin_fmaf(l1r, a1, -b1), f3 = __builtin_fmaf(l3r, a3, b1), f2 = l2 * a2;
        const float p2 = l2 > zero ? f2 : zero;
        const float m2 = l2 <= zero ? f2 : zero;
        float P1 = f3 * l3r, M1 = f1 * l1r;
        float P2 = __builtin_fmaf(f3, vt, p2 * ct), M2 = __builtin_fmaf(f1, vt, m2 * ct);
        unrot<AX>(P1, P2);
        unrot<AX>(M1, M2);
        P[0] = f3; P[1] = P1; P[2] = P2;
        M[0] = f1; M[1] = M1; M[2] = M2;
        return;
    }
    float lE1 = zero, l1 = l1r;
    if (hh1) {
        const float t = ei1 * (ej1 - l1r) * Mf::rcp(ej1 - ei1);
        lE1 = l1r - t;
        l1 = t;
    }
    float lE3 = zero, l3 = l3r;
    if (hh3) {
        lE3 = l3r - ej3 * (l3r - ei3) * Mf::rcp(ej3 - ei3);
        l3 = ej3 * (l3r - ei3) * Mf::rcp((ej3 - ei3) + V);
    }
    rx_upwind_coef<AX>(l1, l2, l3, lE1, lE3, a1, a2, a3, b1, vt, ct, P, M);
}

template <int AX, typename T, bool FAST>
__device__ __forceinline__ void rx_edge_d(const HD<T> &I, const HD<T> &J, T dx, T P[3], T M[3], bool care = true) {
    if constexpr (FAST) rx_edge_fast<AX>(I, J, dx, P, M, care);
    else rx_edge_exact<AX>(I, J, dx, P, M);
}

// The edge's upwP/upwM where the edge is active, zeros elsewhere.  The float32
// path evaluates the edge in every lane and selects (inactive lanes see finite
// floored values, and whatever they compute is discarded): no exec-mask
// branch and no zero-filled register copies around it.
template <int AX, typename T, bool FAST>
__device__ __forceinline__ void edge_or_zero(bool active, const HD<T> &I, const HD<T> &J, T dx, T P[3], T M[3]) {
    const T zero = T(0);
    if constexpr (FAST) {
        rx_edge_d<AX, T, FAST>(I, J, dx, P, M, active);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            P[k] = active ? P[k] : zero;
            M[k] = active ? M[k] : zero;
        }
    } else {
        P[0] = P[1] = P[2] = zero;
        M[0] = M[1] = M[2] = zero;
        if (active) rx_edge_d<AX, T, FAST>(I, J, dx, P, M);
    }
}

// The updated depth of one cell, from scratch (roe_solve_2D for that cell);
// used only when momentum corrections need a neighbour outside the window.
// Inlined once, in a loop over the pending neighbours, rather than called: a
// call would make the whole sweep keep its live registers across the ABI.
// (r, c) differ from lane to lane (each lane checks its own pending neighbour),
// so every load here is addressed per lane.  in and m0 may be moved by sh
// elements (to a strip's first row): cell (r, c) is then element r nx + c - sh.
// PK: in's h, hu, hv, z_b are the packed records at q (moved by sh cells too).
template <typename T, bool FAST, bool PK = false>
__device__ __forceinline__ T h1_at(const F<T> &in, const uint8_t *m0, long sh, T dt, int ny, int nx, T dx, int r,
                                  int c, const T *q = nullptr) {
    const long i = long(r) * nx + c - sh;
    auto at = [&](long j) {
        if constexpr (PK) return hd_load_lane_q<T, FAST>(in, q, m0, j);
        else return hd_load_lane<T, FAST>(in, m0, j);
    };
    const HD<T> me = at(i);
    T acc = T(0), P[3], M[3];
    if (c + 1 < nx) {
        const HD<T> o = at(i + 1);
        if (hd_active(me, o)) { rx_edge_d<0, T, FAST>(me, o, dx, P, M); acc = acc + M[0]; }
    }
    if (c > 0) {
        const HD<T> o = at(i - 1);
        if (hd_active(o, me)) { rx_edge_d<0, T, FAST>(o, me, dx, P, M); acc = acc + P[0]; }
    }
    if (r + 1 < ny) {
        const HD<T> o = at(i + nx);
        if (hd_active(me, o)) { rx_edge_d<1, T, FAST>(me, o, dx, P, M); acc = acc + M[0]; }
    }
    if (r > 0) {
        const HD<T> o = at(i - nx);
        if (hd_active(o, me)) { rx_edge_d<1, T, FAST>(o, me, dx, P, M); acc = acc + P[0]; }
    }
    return me.h - dt * acc / dx;
}

// Hydro values carried down a strip from one row to the next.
template <typename T>
struct HCarry {
    HD<T> cur;        // derived values of the current row
    T pn0, pn1, pn2;  // upwP of the edge above the current row
    T zr_n;           // z + z_b of the row above
    int nod_n;        // nodata bit of the row above
    T h1_n;           // updated depth of the row above (once this wavefront updated it)
};

// The carry at a first row (offset ro): that row derived, and the edge from
// the row above it (offset ro_up) when there is one.
// PK: in's h, hu, hv, z_b are the packed records at q (moved like in).
template <typename T, bool FAST, typename RO, bool PK = false>
__device__ __forceinline__ HCarry<T> hcarry_at(const F<T> &in, const uint8_t *m0, RO ro, RO ro_up, bool has_up,
                                               unsigned ucc, T dx, const T *q = nullptr) {
    const T zero = T(0);
    auto load = [&](RO r) {
        if constexpr (PK) return hd_derive<T, FAST>(hr_load_q<false, false>(in, q, m0, r, ucc, zero));
        else return hd_load<T, FAST>(in, m0, r, ucc);
    };
    HCarry<T> hc;
    hc.cur = load(ro);
    hc.pn0 = hc.pn1 = hc.pn2 = zero;
    hc.zr_n = zero;
    hc.nod_n = 0;
    hc.h1_n = zero;
    if (has_up) {
        const HD<T> up = load(ro_up);
        if (hd_active(up, hc.cur)) {
            T P[3], M[3];
            rx_edge_d<1, T, FAST>(up, hc.cur, dx, P, M);
            hc.pn0 = P[0]; hc.pn1 = P[1]; hc.pn2 = P[2];
        }
        hc.zr_n = up.zr;
        hc.nod_n = up.flags & 2;
    }
    return hc;
}

// The carry at row r0 (absolute rows).
template <typename T, bool FAST>
__device__ __forceinline__ HCarry<T> hcarry_init(const F<T> &in, const uint8_t *m0, int r0, int nx, unsigned ucc,
                                                 T dx) {
    return hcarry_at<T, FAST>(in, m0, long(r0) * nx, long(r0 > 0 ? r0 - 1 : 0) * nx, r0 > 0, ucc, dx);
}

// roe_solve_2D (roe_exner.py:535-612) and momentum_corrections (:14-62) of row
// r, with nxt the derived row r+1; returns the updated h, hu, hv and moves the
// carry to row r+1.  mc: this lane's cell takes the momentum corrections (its
// updated depth is complete here); inE / inW: the east / west neighbour's
// updated depth is held by this wavefront; rfirst: the first row this
// wavefront updated (the row above it is not held).
// PK: in's h, hu, hv, z_b are the packed records at q (moved by sh cells too).
template <typename T, bool FAST, bool PK = false>
__device__ __forceinline__ void hydro_row(HCarry<T> &hc, const HD<T> &nxt, int r, int rfirst, int c, bool mc,
                                          bool inE, bool inW, int ny, int nx, T dx, T dt, T dtdx, const F<T> &in,
                                          const uint8_t *m0, long sh, T &h1, T &hu1, T &hv1, const T *q = nullptr) {
    const T T_ = C<T>(DRY_TOL), zero = T(0);
    const HD<T> &cur = hc.cur;
    const bool hasS = r + 1 < ny;
    // x edge (c | c+1) and y edge (r | r+1)
    const HD<T> east = hd_shdn(cur);
    T Px[3], Mx[3], Py[3], My[3];
    edge_or_zero<0, T, FAST>(c + 1 < nx && hd_active(cur, east), cur, east, dx, Px, Mx);
    const T pw0 = shup(Px[0]), pw1 = shup(Px[1]), pw2 = shup(Px[2]);
    edge_or_zero<1, T, FAST>(hasS && hd_active(cur, nxt), cur, nxt, dx, Py, My);
    T f0 = zero, f1 = zero, f2 = zero;
    if (c + 1 < nx) { f0 = f0 + Mx[0]; f1 = f1 + Mx[1]; f2 = f2 + Mx[2]; }
    if (c > 0) { f0 = f0 + pw0; f1 = f1 + pw1; f2 = f2 + pw2; }
    if (hasS) { f0 = f0 + My[0]; f1 = f1 + My[1]; f2 = f2 + My[2]; }
    if (r > 0) { f0 = f0 + hc.pn0; f1 = f1 + hc.pn1; f2 = f2 + hc.pn2; }
    if constexpr (FAST) {
        h1 = cur.h - dtdx * f0;
        hu1 = cur.hu - dtdx * f1;
        hv1 = cur.hv - dtdx * f2;
    } else {
        h1 = cur.h - dt * f0 / dx;
        hu1 = cur.hu - dt * f1 / dx;
        hv1 = cur.hv - dt * f2 / dx;
    }
    // momentum_corrections, interior cells only
    const T h1E = shdn(h1), h1W = shup(h1);
    const T zrE = east.zr, zrW = shup(cur.zr);
    const int fE = east.flags, fW = shup(cur.flags);
    if (mc && r >= 1 && r <= ny - 2 && c >= 1 && c <= nx - 2) {
        // hu/hv survive only in active cells; the blocking tests (which may
        // need a neighbour's updated depth) run only there
        const bool active = (h1 > T_) && !(cur.flags & 2);
        if (!active) {
            hu1 = zero;
            hv1 = zero;
        } else {
            // blocking by a dry neighbour whose bed is above this water
            // surface; neighbours whose updated depth this wavefront does not
            // hold (outside the lanes, above the first row, the row below)
            // are updated from scratch
            const T eta = h1 + cur.zr;
            const bool inN = r > rfirst;
            bool su = (fE & 2) || (fW & 2);
            bool sv = hc.nod_n || (nxt.flags & 2);
            const bool bE = eta < zrE, bW = eta < zrW, bN = eta < hc.zr_n, bS = eta < nxt.zr;
            if (bE && inE && h1E < T_) su = true;
            if (bW && inW && h1W < T_) su = true;
            if (bN && inN && hc.h1_n < T_) sv = true;
            int far = (bE && !inE ? 1 : 0) | (bW && !inW ? 2 : 0) | (bN && !inN ? 4 : 0) | (bS ? 8 : 0);
            if (su) far &= ~3;
            if (sv) far &= ~12;
            while (far) {
                const int k = __builtin_ctz(far);
                far &= far - 1;
                const int rr = r + (k == 2 ? -1 : (k == 3 ? 1 : 0));
                const int cc = c + (k == 0 ? 1 : (k == 1 ? -1 : 0));
                if (h1_at<T, FAST, PK>(in, m0, sh, dt, ny, nx, dx, rr, cc, q) < T_) {
                    if (k < 2) {
                        su = true;
                        far &= ~3;
                    } else {
                        sv = true;
                        far &= ~12;
                    }
                }
            }
            if (su) hu1 = zero;
            if (sv) hv1 = zero;
        }
    }
    hc.pn0 = Py[0]; hc.pn1 = Py[1]; hc.pn2 = Py[2];
    hc.zr_n = cur.zr;
    hc.nod_n = cur.flags & 2;
    hc.h1_n = h1;
    hc.cur = nxt;
}

// roe_solve_2D + momentum_corrections for the rows of one strip.
template <typename T, bool FAST>
__global__ __launch_bounds__(SW *SWPB) void k_rx_hydro_strip(F<T> in, F<T> out, const uint8_t *__restrict__ m0,
                                                             const T *__restrict__ dtp, int ny, int nx, T dx,
                                                             int momcorr) {
    const StripPos p = strip_pos<SO_WIDE, SR_HYDRO>(ny, nx);
    if (p.r0 >= ny) return;  // whole wavefront
    const T dt = *dtp;
    const T dtdx = dt / dx;
    const int c = p.c;
    const unsigned ucc = unsigned(p.cc);
    const bool mc = momcorr && p.out, inE = p.lane <= SO_WIDE - 1, inW = p.lane >= 2;
    HCarry<T> hc = hcarry_init<T, FAST>(in, m0, p.r0, nx, ucc, dx);
    // rows r+1 and r+2 are in flight while row r is computed
    const int rlast = p.r1 < ny - 1 ? p.r1 : ny - 1;  // last row any edge of the strip reads
    HR<T> q1 = hr_load(in, m0, long(p.r0 + 1 <= rlast ? p.r0 + 1 : rlast) * nx, ucc);
    HR<T> q2 = hr_load(in, m0, long(p.r0 + 2 <= rlast ? p.r0 + 2 : rlast) * nx, ucc);
    wait_vm_all();
    for (int r = p.r0; r < p.r1; ++r) {
        const HD<T> nxt = hd_derive<T, FAST>(q1);
        q1 = q2;
        if (r + 3 <= rlast) q2 = hr_load(in, m0, long(r + 3) * nx, ucc);
        T h1, hu1, hv1;
        hydro_row<T, FAST>(hc, nxt, r, p.r0, c, mc, inE, inW, ny, nx, dx, dt, dtdx, in, m0, 0L, h1, hu1, hv1);
        if (p.out) {
            const long ro = long(r) * nx;
            str(out.h, ro, unsigned(c), h1);
            str(out.hu, ro, unsigned(c), hu1);
            str(out.hv, ro, unsigned(c), hv1);
        }
    }
}

// ---------------------------------------------------------------------------
// G + Exner update (+ SSPRK2 combine)
// ---------------------------------------------------------------------------
template <typename T>
struct ED {          // derived per-cell values for the Exner edge
    T h, hu, hv;     // raw (post-fill) hydro state
    T u, v, s;       // velocities with the Exner floors, sqrt of floored depth
    T zb, zbf;       // raw and floored bed
    T zl;            // low part of the compensated float32 bed (0 when unused)
    T G;
    int flags;       // bit0 wet, bit1 nodata
};

// compute_G (roe_exner.py:85-130) in float32 hardware arithmetic from the
// velocities the bed sweep derives anyway (vel2 = u^2 + v^2 of a wet cell):
// theta = (A/B) vel2 h^(-1/3), theta^1.5 = theta sqrt(theta), and the folded
// weight fw = 8 frac (C/D) / (1 - porosity) in place of gamma1 and the
// porosity divide.  One sediment class (the common case) takes a straight
// path whose three constants stay in scalar registers.
__device__ __forceinline__ float mpm_G_term(float vel2, float icb, float irs, float AB, float thc, float w) {
    using Mf = Mth<float, true>;
    float th = AB * vel2 * icb;
    th = pmax(th, 0.0f);
    const float dth = pmax(th - thc, 0.0f);
    const float g2 = Mf::sqrt(dth) * Mf::rcp(th * Mf::sqrt(th) + 1e-5f);
    return w * irs * g2 * dth;
}

// G is exactly zero where the cell is dry, nodata or still (vel2 == 0 gives
// theta = 0 and so dtheta = 0 in every term), so the terms are evaluated only
// where vel2 != 0 (a NaN vel2 still propagates through them).
__device__ __forceinline__ float mpm_G_fast(float h, float vel2, bool nodata, const SedK &sd) {
    using Mf = Mth<float, true>;
    const float T_ = float(DRY_TOL);
    float G = 0.0f;
    if (vel2 != 0.0f && h > T_ && !nodata) {
        const float icb = Mf::exp2(-0.33333333333333333f * Mf::log2(h));
        const float irs = Mf::rsq(h);
        if (sd.n_sed == 1) {
            G = mpm_G_term(vel2, icb, irs, sd.fAB[0], sd.fthc[0], sd.fw[0]);
        } else {
            for (int k = 0; k < sd.n_sed; ++k) G = G + mpm_G_term(vel2, icb, irs, sd.fAB[k], sd.fthc[k], sd.fw[k]);
        }
    }
    return G;
}

template <typename T>
struct ER {          // raw per-cell inputs of the bed sweep
    T h, hu, hv, zb, G;  // G: the stage input's G (read in constant-G mode only)
    T zl;                // compensated-bed low part (0 when the bed is not compensated)
    int m;
};

template <typename T>
__device__ __forceinline__ ER<T> er_load(const F<T> &hyd, const T *zb_in, const T *zb_lo_in, const T *G_in,
                                         const uint8_t *m0, const pxh_sediments &sd, long ro, unsigned c) {
    ER<T> q;
    q.h = ldr(hyd.h, ro, c); q.hu = ldr(hyd.hu, ro, c); q.hv = ldr(hyd.hv, ro, c);
    q.zb = ldr(zb_in, ro, c);
    q.G = sd.mpm ? T(0) : ldr(G_in, ro, c);
    q.zl = T(0);
    if constexpr (sizeof(T) == 4) {
        if (zb_lo_in) q.zl = ldr(zb_lo_in, ro, c);
    }
    q.m = ldr(m0, ro, c);
    return q;
}

template <typename T>
struct OR {          // previous-step values of one cell (SSPRK2 combine)
    T h, hu, hv, G, zb, zl;
};

template <typename T, typename RO>
__device__ __forceinline__ OR<T> or_load(const F<T> &old, RO ro, unsigned c) {
    OR<T> o;
    o.h = ldr(old.h, ro, c); o.hu = ldr(old.hu, ro, c); o.hv = ldr(old.hv, ro, c);
    o.G = ldr(old.G, ro, c); o.zb = ldr(old.z_b, ro, c);
    o.zl = T(0);
    if constexpr (sizeof(T) == 4) {
        if (old.z_b_lo) o.zl = ldr(old.z_b_lo, ro, c);
    }
    return o;
}

template <typename T>
struct CB {          // combined state 0.5 u^n + 0.5 u^(2) of one cell (h, hu, hv, G)
    T h, hu, hv, G;
};

template <typename T>
__device__ __forceinline__ CB<T> cb_make(const OR<T> &o, T h, T hu, T hv, T G) {
    const T hf = T(0.5);
    CB<T> b;
    b.h = hf * o.h + hf * h;
    b.hu = hf * o.hu + hf * hu;
    b.hv = hf * o.hv + hf * hv;
    b.G = hf * o.G + hf * G;
    return b;
}

template <typename T, bool FAST>
__device__ __forceinline__ ED<T> ed_derive(const ER<T> &q, const SedK &sd) {
    const T T_ = C<T>(DRY_TOL), S_ = C<T>(SED_TOL), zero = T(0);
    ED<T> d;
    d.h = q.h; d.hu = q.hu; d.hv = q.hv;
    const bool nod = q.m != 0;
    const bool wet = d.h > T_;
    const T hf = wet ? d.h : C<T>(DRY_FLOOR);
    if constexpr (FAST) {
        using Mf = Mth<float, true>;
        const float r = Mf::rsq(hf);  // 1 / sqrt(hf): gives 1 / hf and sqrt(hf)
        const float ih = r * r;
        d.u = wet ? d.hu * ih : zero;
        d.v = wet ? d.hv * ih : zero;
        d.s = hf * r;
        d.G = sd.mpm ? mpm_G_fast(d.h, d.u * d.u + d.v * d.v, nod, sd) : const_G<T>(d.h, q.G, nod);
    } else {
        d.G = sd.mpm ? mpm_G<T>(d.h, d.hu, d.hv, nod, sd) : const_G<T>(d.h, q.G, nod);
        d.u = wet ? d.hu / hf : zero;
        d.v = wet ? d.hv / hf : zero;
        d.s = d_sqrt(hf);
    }
    d.zb = q.zb;
    d.zl = q.zl;
    d.zbf = d.zb > S_ ? d.zb : C<T>(SED_FLOOR);
    d.flags = (wet ? 1 : 0) | (nod ? 2 : 0);
    return d;
}

template <typename T, bool FAST>
__device__ __forceinline__ ED<T> ed_load(const F<T> &hyd, const T *zb_in, const T *zb_lo_in, const T *G_in,
                                         const uint8_t *m0, const SedK &sd, long ro, unsigned c) {
    return ed_derive<T, FAST>(er_load(hyd, zb_in, zb_lo_in, G_in, m0, sd, ro, c), sd);
}

template <typename T>
__device__ __forceinline__ ED<T> ed_shdn(const ED<T> &a) {
    ED<T> b;
    b.h = shdn(a.h); b.hu = shdn(a.hu); b.hv = shdn(a.hv);
    b.u = shdn(a.u); b.v = shdn(a.v); b.s = shdn(a.s);
    b.zb = shdn(a.zb); b.zbf = shdn(a.zbf); b.G = shdn(a.G); b.flags = shdn(a.flags);
    return b;
}

template <typename T>
__device__ __forceinline__ bool ed_active(const ED<T> &a, const ED<T> &b) {
    const T S_ = C<T>(SED_TOL);
    const bool stop = ((a.flags | b.flags) & 2) || ((a.zb < S_) && (b.zb < S_));
    return ((a.flags & 1) || (b.flags & 1)) && !stop;
}

// exner_solver (roe_exner.py:614-663) on derived values.
template <int AX, typename T, bool FAST>
__device__ __forceinline__ T ex_edge_d(const ED<T> &I, const ED<T> &J) {
    const T S_ = C<T>(SED_TOL), zero = T(0);
    using Mx = Mth<T, FAST>;
    const T uhi = (AX == 0) ? I.u : -I.v;
    const T uhj = (AX == 0) ? J.u : -J.v;
    const T umi = sq(I.u) + sq(I.v);
    const T umj = sq(J.u) + sq(J.v);
    const T qbi = I.G * umi * uhi;
    const T qbj = J.G * umj * uhj;
    const T gt = T(0.5) * (I.G + J.G);
    const T dz = J.zbf - I.zbf;
    const T dq = gt * umj * uhj - gt * umi * uhi;
    bool up;  // lambda_4 >= 0
    if constexpr (FAST) {
        // the signs of the quotients, without dividing: dz is nonzero in the
        // first form (|dz| > SED_TOL), s_i + s_j > 0 in the second; a zero
        // numerator gives lambda_4 = 0 >= 0 and a NaN one lambda_4 = NaN < 0
        const float num = uhi * I.s + uhj * J.s;
        up = d_abs(dz) > S_ ? (dq == zero || (dz > zero ? dq > zero : dq < zero)) : num >= zero;
    } else if (d_abs(dz) > S_) {
        up = Mx::div(dq, dz) >= zero;
    } else {
        up = Mx::div(uhi * I.s + uhj * J.s, I.s + J.s) >= zero;
    }
    const T ci = (gt - I.G) * umi * uhi;
    const T cj = (gt - J.G) * umj * uhj;
    return up ? (qbi + ci) : (qbj + cj);
}

// Compensated float32 bed: (hi, lo) - inc by TwoSum and Fast2Sum.
__device__ __forceinline__ void cb_sub(float hi, float lo, float inc, float &ohi, float &olo) {
    const float b = -inc;
    const float s = hi + b;
    const float bb = s - hi;
    const float e = (hi - (s - bb)) + (b - bb);
    const float l = lo + e;
    const float h2 = s + l;
    olo = l - (h2 - s);
    ohi = h2;
}
// 0.5 (aH + aL) + 0.5 (bH + bL), renormalised.
__device__ __forceinline__ void cb_half_sum(float aH, float aL, float bH, float bL, float &ohi, float &olo) {
    const float x = 0.5f * aH, y = 0.5f * bH;
    const float s = x + y;
    const float bb = s - x;
    const float e = (x - (s - bb)) + (y - bb);
    const float l = e + (0.5f * aL + 0.5f * bL);
    const float h2 = s + l;
    olo = l - (h2 - s);
    ohi = h2;
}

// ---------------------------------------------------------------------------
// CFL speed per edge (roe_exner.py:211-255) on per-cell values
// ---------------------------------------------------------------------------
template <typename T>
struct DD {          // per-cell values for the dt edge
    T h;             // raw depth (active test)
    T hf, u, v, s;   // max(h, DRY_TOL), hu/hf, hv/hf (no wet guard, as coded), sqrt(hf)
    T G;
    int nod;
};

template <typename T, bool FAST>
__device__ __forceinline__ DD<T> dd_make(T h, T hu, T hv, T G, int nod) {
    DD<T> d;
    d.h = h;
    d.hf = pmax(h, C<T>(DRY_TOL));
    if constexpr (FAST) {
        using Mf = Mth<float, true>;
        const float r = Mf::rsq(d.hf);  // 1 / sqrt(hf): gives 1 / hf and sqrt(hf)
        const float ih = r * r;
        d.u = hu * ih;
        d.v = hv * ih;
        d.s = d.hf * r;
    } else {
        d.u = hu / d.hf;
        d.v = hv / d.hf;
        d.s = d_sqrt(d.hf);
    }
    d.G = G;
    d.nod = nod;
    return d;
}

template <typename T>
__device__ __forceinline__ DD<T> dd_shdn(const DD<T> &a) {
    DD<T> b;
    b.h = shdn(a.h); b.hf = shdn(a.hf); b.u = shdn(a.u); b.v = shdn(a.v); b.s = shdn(a.s);
    b.G = shdn(a.G); b.nod = shdn(a.nod);
    return b;
}

template <typename T>
__device__ __forceinline__ bool dd_active(const DD<T> &a, const DD<T> &b) {
    const T T_ = C<T>(DRY_TOL);
    return ((a.h > T_) || (b.h > T_)) && !(a.nod || b.nod);
}

// get_lambda in float32 hardware arithmetic (float32 FAST path only).  The
// coded cubic (a = -1, b = -2 ut, c = (1 + at) ct^2 - ut^2, d = -ct^2 at ut)
// has the depressed form p = -((1 + at) ct^2 + ut^2 / 3) < 0,
// q = ut (ct^2 (18 + 45 at) - 2 ut^2) / 27 and shift -b / (3 a) = -2 ut / 3,
// so with P = -p: x = -(3 sqrt(3) / 2) q / P^(3/2) and the scale
// 2 sqrt(-p / 3) = (2 / sqrt(3)) sqrt(P).  The largest root of
// 4 t^3 - 3 t = |x| takes two Newton steps from the quadratic fit (error
// below 2e-3, quadratic convergence with factor <= 2, so < 1e-10 after two);
// the other extreme root follows by deflation, as in trisect.  Next to
// |x| = 1 the root pair is nearly double and its roots go with the square
// root of the error in x, so there (|x| >= 0.999, and the clipped |x| >= 1)
// it returns false and the caller evaluates the double form; below, max|lambda|
// stays within 3e-6 of the double form.
__device__ __forceinline__ bool get_lambda_f32(float ut, float at, float ct, float &lmin, float &lmax) {
#pragma clang fp contract(on)
    using Mf = Mth<float, true>;
    const float ct2 = ct * ct;
    const float P = (1.0f + at) * ct2 + ut * ut * 0.33333333333333333f;
    const float q = ut * (ct2 * (18.0f + 45.0f * at) - 2.0f * ut * ut) * 0.037037037037037035f;
    const float rP = Mf::rsq(P);
    const float x = -2.598076211353316f * q * (rP * rP * rP);
    const float st = 1.1547005383792515f * P * rP;
    const float sh = -0.66666666666666667f * ut;
    const float xa = __builtin_fabsf(x);
    if (!(xa < 0.999f)) return false;
    float t = 0.8660254037844386f + xa * (0.16666666666666666f - 0.03269207045111198f * xa);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float t2 = t * t;
        t = t - ((4.0f * t2 - 3.0f) * t - xa) * Mf::rcp(12.0f * t2 - 3.0f);
    }
    const float other = 0.5f * (-t - Mf::sqrt(3.0f * (1.0f - t) * (1.0f + t)));
    float tmin, tmax;
    if (x < 0.0f) {
        tmin = -t;
        tmax = -other;
    } else {
        tmax = t;
        tmin = other;
    }
    lmax = st * tmax + sh;
    lmin = st * tmin + sh;
    return true;
}

template <int AX, typename T, bool FAST>
__device__ __forceinline__ T dt_edge_d(const DD<T> &I, const DD<T> &J, T dx) {
    const T g = C<T>(GRAV);
    T uhi, vhi, uhj, vhj;
    rot<AX>(I.u, I.v, uhi, vhi);
    rot<AX>(J.u, J.v, uhj, vhj);
    const T ht = T(0.5) * (I.hf + J.hf);
    const T gt = T(0.5) * (I.G + J.G);
    const T S = sq(uhi) + uhi * uhj + sq(uhj) + vhi * vhj;
    T ut, ct, at;
    if constexpr (FAST) {
        using Mf = Mth<float, true>;
        const float sum = I.s + J.s, prod = I.s * J.s;
        const float R = Mf::rcp(sum * prod);  // 1 / (s_i + s_j) = R s_i s_j, 1 / (s_i s_j) = R (s_i + s_j)
        ut = (uhi * I.s + uhj * J.s) * (R * prod);
        ct = Mf::sqrt(g * ht);
        at = gt * S * (R * sum);
    } else {
        ut = (uhi * I.s + uhj * J.s) / (I.s + J.s);
        ct = d_sqrt(g * ht);
        at = gt * S / (I.s * J.s);
    }
    const T l1 = ut - ct, l3 = ut + ct;
    const T ss = pmax(d_abs(l1), d_abs(l3));
    T rs;
    if (at < C<T>(1e-5)) {
        rs = ss;  // linear branch of _get_lambda: (ut - ct, ut + ct)
    } else if constexpr (FAST) {
        float lmin, lmax;
        if (get_lambda_f32(ut, at, ct, lmin, lmax)) {
            rs = pmax(d_abs(lmin), d_abs(lmax));
        } else {
            double dmin, dmax;
            get_lambda(static_cast<double>(ut), static_cast<double>(at), static_cast<double>(ct), dmin, dmax);
            rs = pmax(d_abs(static_cast<float>(dmin)), d_abs(static_cast<float>(dmax)));
        }
    } else {
        double lmin, lmax;
        get_lambda(static_cast<double>(ut), static_cast<double>(at), static_cast<double>(ct), lmin, lmax);
        rs = pmax(d_abs(static_cast<T>(lmin)), d_abs(static_cast<T>(lmax)));
    }
    const T sp = (rs < C<T>(VEL_TOL)) ? ss : rs;
    return Mth<T, FAST>::div(dx, sp);
}

// Minimum over a block of NW wavefronts; the result is valid in thread 0.
template <int NW, typename T>
__device__ __forceinline__ T wave_block_min_n(T v, T *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & (SW - 1)) == 0) red[threadIdx.x / SW] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < NW; ++k) v = pmin(v, red[k]);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_block_min(T v, T *red) {
    return wave_block_min_n<SWPB>(v, red);
}

// MODE 0: out.z_b = updated bed (+ out.G when write_G).
// MODE 1: out = 0.5 * old + 0.5 * (stage result) over h, hu, hv, z_b, G; with
// DT also the CFL dt partial minimum of the combined state into partial[block].
// zb_lo_in / out.z_b_lo / old.z_b_lo: compensated bed (float32), else null.
// Every per-row input (the stage rows and the previous-step rows) is loaded
// two rows ahead of its use.
template <typename T, bool FAST, int MODE, bool DT>
__global__ __launch_bounds__(SW *SWPB) void k_rx_bed_strip(F<T> hyd, const T *__restrict__ zb_in,
                                                           const T *__restrict__ zb_lo_in, const T *__restrict__ G_in,
                                                           F<T> out, F<T> old, const uint8_t *__restrict__ m0,
                                                           const T *__restrict__ dtp, int ny, int nx, T dx,
                                                           SedK sd, int write_G, T *__restrict__ partial,
                                                           T *__restrict__ xs) {
    __shared__ T red[SWPB];
    const StripPos p = strip_pos<SO_WIDE, SR_BED>(ny, nx);
    const T zero = T(0), hf = T(0.5);
    const int c = p.c;
    T vmin = inf_v<T>();
    if (p.r0 < ny) {
        const T dt = *dtp;
        const T dtdx = dt / dx;
        const int rlast = p.r1 < ny - 1 ? p.r1 : ny - 1;
        const unsigned ucc = unsigned(p.cc), uc = unsigned(c);
        auto row = [&](int r) { return long(r <= rlast ? r : rlast) * nx; };
        ED<T> cur = ed_load<T, FAST>(hyd, zb_in, zb_lo_in, G_in, m0, sd, row(p.r0), ucc);
        T fn = zero;  // flux through the edge (r-1 | r)
        if (p.r0 > 0) {
            const ED<T> up = ed_load<T, FAST>(hyd, zb_in, zb_lo_in, G_in, m0, sd, long(p.r0 - 1) * nx, ucc);
            if (ed_active(up, cur)) fn = ex_edge_d<1, T, FAST>(up, cur);
        }
        ER<T> q1 = er_load(hyd, zb_in, zb_lo_in, G_in, m0, sd, row(p.r0 + 1), ucc);
        ER<T> q2 = er_load(hyd, zb_in, zb_lo_in, G_in, m0, sd, row(p.r0 + 2), ucc);
        // combine: the combined state and previous-step bed of row r; the
        // previous-step rows r+1 and r+2 in flight
        CB<T> cbc = {zero, zero, zero, zero};
        T ozb = zero, ozl = zero;
        OR<T> o1 = {zero, zero, zero, zero, zero, zero}, o2 = o1;
        if (MODE == 1) {
            const OR<T> o0 = or_load(old, row(p.r0), ucc);
            cbc = cb_make(o0, cur.h, cur.hu, cur.hv, cur.G);
            ozb = o0.zb;
            ozl = o0.zl;
            o1 = or_load(old, row(p.r0 + 1), ucc);
            o2 = or_load(old, row(p.r0 + 2), ucc);
        }
        DD<T> dcur;
        if (DT) dcur = dd_make<T, FAST>(cbc.h, cbc.hu, cbc.hv, cbc.G, cur.flags & 2);
        wait_vm_all();
        for (int r = p.r0; r < p.r1; ++r) {
            const bool hasS = r + 1 < ny;
            const ED<T> nxt = ed_derive<T, FAST>(q1, sd);
            // Row r+1's previous-step values are consumed here, before the
            // refill: a register set still read after its refill is issued
            // would be copied at the loop latch, waiting for the new loads.
            CB<T> cbn = cbc;
            T ozb_n = zero, ozl_n = zero;
            if (MODE == 1) {
                cbn = cb_make(o1, nxt.h, nxt.hu, nxt.hv, nxt.G);
                ozb_n = o1.zb;
                ozl_n = o1.zl;
                o1 = o2;
            }
            q1 = q2;
            if (r + 3 <= rlast) {
                q2 = er_load(hyd, zb_in, zb_lo_in, G_in, m0, sd, long(r + 3) * nx, ucc);
                if (MODE == 1) o2 = or_load(old, long(r + 3) * nx, ucc);
            }
            const ED<T> east = ed_shdn(cur);
            T fx = zero;
            if (c + 1 < nx && ed_active(cur, east)) fx = ex_edge_d<0, T, FAST>(cur, east);
            const T fw = shup(fx);
            T fy = zero;
            if (hasS && ed_active(cur, nxt)) fy = ex_edge_d<1, T, FAST>(cur, nxt);
            T div = zero;
            if (c > 0) div = div + (-fw);
            if (c + 1 < nx) div = div + fx;
            if (r > 0) div = div + (-fn);
            if (hasS) div = div + fy;
            if (p.out) {
                const long i = long(r) * nx;  // row offset; column uc
                const T inc = FAST ? dtdx * div : dt * div / dx;
                T zb1 = cur.zb - inc, zl1 = zero;
                if constexpr (sizeof(T) == 4) {
                    if (zb_lo_in) cb_sub(cur.zb, cur.zl, inc, zb1, zl1);
                }
                if (MODE == 0) {
                    str(out.z_b, i, uc, zb1);
                    if constexpr (sizeof(T) == 4) {
                        if (out.z_b_lo) str(out.z_b_lo, i, uc, zl1);
                    }
                    if (write_G) str(out.G, i, uc, cur.G);
                } else {
                    str(out.h, i, uc, cbc.h);
                    str(out.hu, i, uc, cbc.hu);
                    str(out.hv, i, uc, cbc.hv);
                    str(out.G, i, uc, cbc.G);
                    T zc = hf * ozb + hf * zb1;
                    if constexpr (sizeof(T) == 4) {
                        if (out.z_b_lo) {
                            T zcl;
                            cb_half_sum(ozb, ozl, zb1, zl1, zc, zcl);
                            str(out.z_b_lo, i, uc, zcl);
                        }
                    }
                    str(out.z_b, i, uc, zc);
                }
            }
            if (MODE == 1) {
                if (DT) {
                    // edges (c | c+1) of row r and (r | r+1), owned by the output
                    // lanes, only inside the strip: the combined value of row r1 and
                    // of the column right of the last output lane belongs to a
                    // neighbouring strip, which may already have stored it over
                    // `old` (k_rx_seam_dt takes those edges)
                    if (p.out && (p.lane == 1 || p.lane == SO_WIDE))
                        seam_put(seam_rec(xs, (p.c + 1 - p.lane) / SO_WIDE, r, ny), p.lane != 1, cbc.h, cbc.hu,
                                 cbc.hv, cbc.G, (cur.flags & 2) != 0);
                    DD<T> dnxt = dcur;
                    if (hasS) dnxt = dd_make<T, FAST>(cbn.h, cbn.hu, cbn.hv, cbn.G, nxt.flags & 2);
                    const DD<T> deast = dd_shdn(dcur);
                    if (p.out && p.lane < SO_WIDE && c + 1 < nx && dd_active(dcur, deast))
                        vmin = pmin(vmin, dt_edge_d<0, T, FAST>(dcur, deast, dx));
                    if (p.out && r + 1 < p.r1 && dd_active(dcur, dnxt))
                        vmin = pmin(vmin, dt_edge_d<1, T, FAST>(dcur, dnxt, dx));
                    dcur = dnxt;
                }
                cbc = cbn;
                ozb = ozb_n;
                ozl = ozl_n;
            }
            fn = fy;
            cur = nxt;
        }
    }
    if (DT) {
        vmin = wave_block_min(vmin, red);
        if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = vmin;
    }
}

// ---------------------------------------------------------------------------
// CFL dt partial minima (roe_exner.py:257-310), one partial per block
// ---------------------------------------------------------------------------
template <typename T>
struct DR {          // raw per-cell inputs of the dt sweep
    T h, hu, hv, G;
    int m;
};

template <typename T>
__device__ __forceinline__ DR<T> dr_load(const F<T> &s, const uint8_t *m0, long ro, unsigned c) {
    DR<T> q;
    q.h = ldr(s.h, ro, c); q.hu = ldr(s.hu, ro, c); q.hv = ldr(s.hv, ro, c); q.G = ldr(s.G, ro, c);
    q.m = ldr(m0, ro, c);
    return q;
}

template <typename T, bool FAST>
__device__ __forceinline__ DD<T> dd_from(const DR<T> &q) {
    return dd_make<T, FAST>(q.h, q.hu, q.hv, q.G, q.m);
}

template <typename T, bool FAST>
__device__ __forceinline__ T dt_strip_run(const StripPos p, const F<T> s, const uint8_t *__restrict__ m0, int ny,
                                          int nx, T dx) {
    T v = inf_v<T>();
    if (p.r0 < ny) {
        const unsigned ucc = unsigned(p.cc);
        const int rlast = p.r1 < ny - 1 ? p.r1 : ny - 1;
        auto row = [&](int r) { return long(r <= rlast ? r : rlast) * nx; };
        DD<T> cur = dd_from<T, FAST>(dr_load(s, m0, row(p.r0), ucc));
        // rows r+1 and r+2 in flight while row r is evaluated
        DR<T> q1 = dr_load(s, m0, row(p.r0 + 1), ucc);
        DR<T> q2 = dr_load(s, m0, row(p.r0 + 2), ucc);
        wait_vm_all();
        for (int r = p.r0; r < p.r1; ++r) {
            const bool hasS = r + 1 < ny;
            DD<T> nxt = cur;
            if (hasS) nxt = dd_from<T, FAST>(q1);
            q1 = q2;
            if (r + 3 <= rlast) q2 = dr_load(s, m0, long(r + 3) * nx, ucc);
            const DD<T> east = dd_shdn(cur);
            // lanes 1..62 own the x edge (c | c+1) and the y edge (r | r+1) of their column
            if (p.out && p.c + 1 < nx && dd_active(cur, east)) v = pmin(v, dt_edge_d<0, T, FAST>(cur, east, dx));
            if (p.out && hasS && dd_active(cur, nxt)) v = pmin(v, dt_edge_d<1, T, FAST>(cur, nxt, dx));
            cur = nxt;
        }
    }
    return v;
}

template <typename T, bool FAST>
__global__ __launch_bounds__(SW *SWPB) void k_rx_dt_strip(F<T> s, const uint8_t *__restrict__ m0, int ny, int nx,
                                                          T dx, T *__restrict__ partial) {
    __shared__ T red[SWPB];
    const StripPos p = strip_pos<SO_WIDE, SR_BED>(ny, nx);
    T v = dt_strip_run<T, FAST>(p, s, m0, ny, nx, dx);
    v = wave_block_min(v, red);
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = v;
}

}  // namespace pxh
