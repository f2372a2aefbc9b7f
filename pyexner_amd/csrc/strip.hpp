// Wavefront-strip sweeps for the Roe-Exner stage.
//
// A 64-lane wavefront owns a strip of 62 output columns (lanes 1..62; lanes 0
// and 63 are the west/east halo columns) and walks down SROWS rows keeping
// the current and next row in registers.  Every edge flux is evaluated once:
// the x edge (c | c+1) by lane c with lane c+1's values from a cross-lane
// shuffle (its upwP reaches lane c+1 by the reverse shuffle), the y edge
// (r | r+1) by the lane that owns the column, carried to the next row in a
// register.  Row loads are coalesced; there is no integer division.
//
// FAST selects the float32 arithmetic: hardware reciprocal, reciprocal square
// root, square root, log2 and exp2 (about 1 ulp each) with reciprocals shared
// across an edge.  FAST = false keeps the reference's operation sequence with
// IEEE division and square root (the double build, and float32 when exact
// agreement with a float32 restatement is wanted).
#pragma once
#include "common.hpp"
#include "flux_roe_exner.hpp"

namespace pxh {

constexpr int SW = 64;       // lanes per strip
constexpr int SOUT = 62;     // output columns per strip
constexpr int SROWS = 32;    // rows per strip
constexpr int SWPB = 4;      // strips (wavefronts) per block, stacked in y

struct StripPos {
    int lane, c, cc, r0, r1;
    bool out;
};

__device__ __forceinline__ StripPos strip_pos(int ny, int nx) {
    StripPos p;
    p.lane = threadIdx.x & (SW - 1);
    const int w = threadIdx.x / SW;
    p.c = blockIdx.x * SOUT - 1 + p.lane;
    p.cc = p.c < 0 ? 0 : (p.c >= nx ? nx - 1 : p.c);
    p.out = p.lane >= 1 && p.lane <= SOUT && p.c < nx;
    p.r0 = (blockIdx.y * SWPB + w) * SROWS;
    p.r1 = p.r0 + SROWS < ny ? p.r0 + SROWS : ny;
    return p;
}

static inline dim3 strip_grid(int ny, int nx) {
    const int sx = (nx + SOUT - 1) / SOUT;
    const int sy = (ny + SROWS - 1) / SROWS;
    return dim3(sx, (sy + SWPB - 1) / SWPB);
}

template <typename T>
__device__ __forceinline__ T shdn(T v) { return __shfl_down(v, 1, SW); }
template <typename T>
__device__ __forceinline__ T shup(T v) { return __shfl_up(v, 1, SW); }

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
    int flags;       // bit0 wet (h > DRY_TOL), bit1 nodata
};

template <typename T, bool FAST>
__device__ __forceinline__ HD<T> hd_load(const F<T> &s, const uint8_t *m0, long i) {
    const T T_ = C<T>(DRY_TOL), S_ = C<T>(SED_TOL), g = C<T>(GRAV), zero = T(0);
    HD<T> d;
    d.h = s.h[i]; d.hu = s.hu[i]; d.hv = s.hv[i];
    const T zb = s.z_b[i], z = s.z[i];
    d.n = s.n_b[i];
    d.hf = pmax(d.h, T_);
    const bool wet = d.hf > T_;
    if constexpr (FAST) {
        using M = Mth<float, true>;
        const float ih = M::rcp(d.hf);
        d.u = wet ? d.hu * ih : zero;
        d.v = wet ? d.hv * ih : zero;
        d.s = M::sqrt(d.hf);
        d.e = d.s * 3.1320919526731650f;  // sqrt(9.81)
    } else {
        d.u = wet ? d.hu / d.hf : zero;
        d.v = wet ? d.hv / d.hf : zero;
        d.s = d_sqrt(d.hf);
        d.e = d_sqrt(g * d.hf);
    }
    d.zt = z + (zb > S_ ? zb : C<T>(SED_FLOOR));
    d.zr = z + zb;
    d.flags = (d.h > T_ ? 1 : 0) | (m0[i] ? 2 : 0);
    return d;
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

// The same edge in float32 hardware arithmetic: one reciprocal for the Roe
// average, one reciprocal square root for the celerity and its inverse, and
// h~^(4/3) as exp2(4/3 log2 h~).
template <int AX>
__device__ __forceinline__ void rx_edge_fast(const HD<float> &I, const HD<float> &J, float dx, float P[3],
                                             float M[3]) {
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
    float ei = uhi - I.e, ej = uhj - J.e;
    float lE1 = zero, l1 = l1r;
    if ((ei < zero) && (ej > zero)) {
        const float t = ei * (ej - l1r) * Mf::rcp(ej - ei);
        lE1 = l1r - t;
        l1 = t;
    }
    ei = uhi + I.e;
    ej = uhj + J.e;
    float lE3 = zero, l3 = l3r;
    if ((ei < zero) && (ej > zero)) {
        lE3 = l3r - ej * (l3r - ei) * Mf::rcp(ej - ei);
        l3 = ej * (l3r - ei) * Mf::rcp((ej - ei) + V);
    }
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
        const float h43 = Mf::exp2(1.3333333333333333f * Mf::log2(ht));
        const float sf = (nt * nt * Mf::sqrt(ut * ut + vt * vt) * ut) * Mf::rcp(pmax(T_, h43));
        src = src - g * ht * sf * dx;
    }
    const float b1 = src * ih2;
    rx_upwind<AX>(l1, l2, l3, lE1, lE3, a1, a2, a3, b1, vt, ct, P, M);
}

template <int AX, typename T, bool FAST>
__device__ __forceinline__ void rx_edge_d(const HD<T> &I, const HD<T> &J, T dx, T P[3], T M[3]) {
    if constexpr (FAST) rx_edge_fast<AX>(I, J, dx, P, M);
    else rx_edge_exact<AX>(I, J, dx, P, M);
}

// The updated depth of one cell, from scratch (roe_solve_2D for that cell);
// used only when momentum corrections need a neighbour outside the window.
template <typename T, bool FAST>
__device__ __noinline__ T h1_at(const F<T> &in, const uint8_t *m0, T dt, int ny, int nx, T dx, int r, int c) {
    const long i = long(r) * nx + c;
    const HD<T> me = hd_load<T, FAST>(in, m0, i);
    T acc = T(0), P[3], M[3];
    if (c + 1 < nx) {
        const HD<T> o = hd_load<T, FAST>(in, m0, i + 1);
        if (hd_active(me, o)) { rx_edge_d<0, T, FAST>(me, o, dx, P, M); acc = acc + M[0]; }
    }
    if (c > 0) {
        const HD<T> o = hd_load<T, FAST>(in, m0, i - 1);
        if (hd_active(o, me)) { rx_edge_d<0, T, FAST>(o, me, dx, P, M); acc = acc + P[0]; }
    }
    if (r + 1 < ny) {
        const HD<T> o = hd_load<T, FAST>(in, m0, i + nx);
        if (hd_active(me, o)) { rx_edge_d<1, T, FAST>(me, o, dx, P, M); acc = acc + M[0]; }
    }
    if (r > 0) {
        const HD<T> o = hd_load<T, FAST>(in, m0, i - nx);
        if (hd_active(o, me)) { rx_edge_d<1, T, FAST>(o, me, dx, P, M); acc = acc + P[0]; }
    }
    return me.h - dt * acc / dx;
}

// roe_solve_2D + momentum_corrections for the rows of one strip.
template <typename T, bool FAST>
__global__ __launch_bounds__(SW *SWPB) void k_rx_hydro_strip(F<T> in, F<T> out, const uint8_t *__restrict__ m0,
                                                             const T *__restrict__ dtp, int ny, int nx, T dx,
                                                             int momcorr) {
    const StripPos p = strip_pos(ny, nx);
    if (p.r0 >= ny) return;  // whole wavefront
    const T T_ = C<T>(DRY_TOL), zero = T(0);
    const T dt = *dtp;
    const T dtdx = dt / dx;
    const int c = p.c;
    HD<T> cur = hd_load<T, FAST>(in, m0, long(p.r0) * nx + p.cc);
    T pn0 = zero, pn1 = zero, pn2 = zero;  // upwP of the edge (r-1 | r)
    T zr_n = zero;                          // z + z_b of row r-1
    int nod_n = 0;
    if (p.r0 > 0) {
        const HD<T> up = hd_load<T, FAST>(in, m0, long(p.r0 - 1) * nx + p.cc);
        if (hd_active(up, cur)) {
            T P[3], M[3];
            rx_edge_d<1, T, FAST>(up, cur, dx, P, M);
            pn0 = P[0]; pn1 = P[1]; pn2 = P[2];
        }
        zr_n = up.zr;
        nod_n = up.flags & 2;
    }
    T h1_n = zero;  // updated depth of row r-1 (valid from the second row on)
    for (int r = p.r0; r < p.r1; ++r) {
        const bool hasS = r + 1 < ny;
        HD<T> nxt = cur;
        if (hasS) nxt = hd_load<T, FAST>(in, m0, long(r + 1) * nx + p.cc);
        // x edge (c | c+1)
        const HD<T> east = hd_shdn(cur);
        T Px[3] = {zero, zero, zero}, Mx[3] = {zero, zero, zero};
        if (c + 1 < nx && hd_active(cur, east)) rx_edge_d<0, T, FAST>(cur, east, dx, Px, Mx);
        const T pw0 = shup(Px[0]), pw1 = shup(Px[1]), pw2 = shup(Px[2]);
        // y edge (r | r+1)
        T Py[3] = {zero, zero, zero}, My[3] = {zero, zero, zero};
        if (hasS && hd_active(cur, nxt)) rx_edge_d<1, T, FAST>(cur, nxt, dx, Py, My);
        T f0 = zero, f1 = zero, f2 = zero;
        if (c + 1 < nx) { f0 = f0 + Mx[0]; f1 = f1 + Mx[1]; f2 = f2 + Mx[2]; }
        if (c > 0) { f0 = f0 + pw0; f1 = f1 + pw1; f2 = f2 + pw2; }
        if (hasS) { f0 = f0 + My[0]; f1 = f1 + My[1]; f2 = f2 + My[2]; }
        if (r > 0) { f0 = f0 + pn0; f1 = f1 + pn1; f2 = f2 + pn2; }
        T h1, hu1, hv1;
        if constexpr (FAST) {
            h1 = cur.h - dtdx * f0;
            hu1 = cur.hu - dtdx * f1;
            hv1 = cur.hv - dtdx * f2;
        } else {
            h1 = cur.h - dt * f0 / dx;
            hu1 = cur.hu - dt * f1 / dx;
            hv1 = cur.hv - dt * f2 / dx;
        }
        // momentum_corrections (roe_exner.py:14-62), interior cells only
        const T h1E = shdn(h1), h1W = shup(h1);
        const T zrE = shdn(cur.zr), zrW = shup(cur.zr);
        const int fE = shdn(cur.flags), fW = shup(cur.flags);
        if (momcorr && p.out && r >= 1 && r <= ny - 2 && c >= 1 && c <= nx - 2) {
            // hu/hv survive only in active cells; the blocking tests (which may
            // need a neighbour's updated depth) run only there
            const bool active = (h1 > T_) && !(cur.flags & 2);
            if (!active) {
                hu1 = zero;
                hv1 = zero;
            } else {
                const T eta = h1 + cur.zr;
                bool su = (fE & 2) || (fW & 2);
                if (!su && eta < zrE) {
                    const T hE = (p.lane <= SOUT - 1) ? h1E : h1_at<T, FAST>(in, m0, dt, ny, nx, dx, r, c + 1);
                    su = hE < T_;
                }
                if (!su && eta < zrW) {
                    const T hW = (p.lane >= 2) ? h1W : h1_at<T, FAST>(in, m0, dt, ny, nx, dx, r, c - 1);
                    su = hW < T_;
                }
                bool sv = nod_n || (nxt.flags & 2);
                if (!sv && eta < zr_n) {
                    const T hN = (r > p.r0) ? h1_n : h1_at<T, FAST>(in, m0, dt, ny, nx, dx, r - 1, c);
                    sv = hN < T_;
                }
                if (!sv && eta < nxt.zr) sv = h1_at<T, FAST>(in, m0, dt, ny, nx, dx, r + 1, c) < T_;
                if (su) hu1 = zero;
                if (sv) hv1 = zero;
            }
        }
        if (p.out) {
            const long i = long(r) * nx + c;
            out.h[i] = h1;
            out.hu[i] = hu1;
            out.hv[i] = hv1;
        }
        pn0 = Py[0]; pn1 = Py[1]; pn2 = Py[2];
        zr_n = cur.zr;
        nod_n = cur.flags & 2;
        h1_n = h1;
        cur = nxt;
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
    T G;
    int flags;       // bit0 wet, bit1 nodata
};

// compute_G (roe_exner.py:85-130) in float32 hardware arithmetic:
// theta = (A/B) vel2 h^(-1/3), gamma1 = (C/D) h^(-1/2), theta^1.5 = theta sqrt(theta).
__device__ __forceinline__ float mpm_G_fast(float h, float hu, float hv, bool nodata, const pxh_sediments &sd) {
    using Mf = Mth<float, true>;
    const float T_ = float(DRY_TOL), zero = 0.0f;
    h = pmax(h, T_);
    const float ih = Mf::rcp(h);
    const float u = hu * ih, v = hv * ih;
    const float vel2 = u * u + v * v;
    const float icb = Mf::exp2(-0.33333333333333333f * Mf::log2(h));
    const float irs = Mf::rsq(h);
    float G = zero;
    for (int k = 0; k < sd.n_sed; ++k) {
        float th = float(sd.AB[k]) * vel2 * icb;
        th = pmax(th, zero);
        const float dth = pmax(th - float(sd.theta_c[k]), zero);
        const float g1 = float(sd.CD[k]) * irs;
        const float g2 = 8.0f * Mf::sqrt(dth) * Mf::rcp(th * Mf::sqrt(th) + 1e-5f);
        G = G + float(sd.frac[k]) * g1 * g2 * dth;
    }
    G = G * Mf::rcp(float(sd.one_minus_por));
    if (nodata || !(h > T_)) G = zero;
    return G;
}

template <typename T, bool FAST>
__device__ __forceinline__ ED<T> ed_load(const F<T> &hyd, const T *zb_in, const T *G_in, const uint8_t *m0,
                                         const pxh_sediments &sd, long i) {
    const T T_ = C<T>(DRY_TOL), S_ = C<T>(SED_TOL), zero = T(0);
    ED<T> d;
    d.h = hyd.h[i]; d.hu = hyd.hu[i]; d.hv = hyd.hv[i];
    const bool nod = m0[i] != 0;
    const bool wet = d.h > T_;
    const T hf = wet ? d.h : C<T>(DRY_FLOOR);
    if constexpr (FAST) {
        using Mf = Mth<float, true>;
        d.G = sd.mpm ? mpm_G_fast(d.h, d.hu, d.hv, nod, sd) : const_G<T>(d.h, G_in[i], nod);
        const float ih = Mf::rcp(hf);
        d.u = wet ? d.hu * ih : zero;
        d.v = wet ? d.hv * ih : zero;
        d.s = Mf::sqrt(hf);
    } else {
        d.G = sd.mpm ? mpm_G<T>(d.h, d.hu, d.hv, nod, sd) : const_G<T>(d.h, G_in[i], nod);
        d.u = wet ? d.hu / hf : zero;
        d.v = wet ? d.hv / hf : zero;
        d.s = d_sqrt(hf);
    }
    d.zb = zb_in[i];
    d.zbf = d.zb > S_ ? d.zb : C<T>(SED_FLOOR);
    d.flags = (wet ? 1 : 0) | (nod ? 2 : 0);
    return d;
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
    if (d_abs(dz) > S_) {
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

// MODE 0: out.z_b = updated bed (+ out.G when write_G).
// MODE 1: out = 0.5 * old + 0.5 * (stage result) over h, hu, hv, z_b, G.
// zb_lo_in / out.z_b_lo / old.z_b_lo: compensated bed (float32), else null.
template <typename T, bool FAST, int MODE>
__global__ __launch_bounds__(SW *SWPB) void k_rx_bed_strip(F<T> hyd, const T *__restrict__ zb_in,
                                                           const T *__restrict__ zb_lo_in, const T *__restrict__ G_in,
                                                           F<T> out, F<T> old, const uint8_t *__restrict__ m0,
                                                           const T *__restrict__ dtp, int ny, int nx, T dx,
                                                           pxh_sediments sd, int write_G) {
    const StripPos p = strip_pos(ny, nx);
    if (p.r0 >= ny) return;
    const T zero = T(0);
    const T dt = *dtp;
    const T dtdx = dt / dx;
    const int c = p.c;
    ED<T> cur = ed_load<T, FAST>(hyd, zb_in, G_in, m0, sd, long(p.r0) * nx + p.cc);
    T fn = zero;  // flux through the edge (r-1 | r)
    if (p.r0 > 0) {
        const ED<T> up = ed_load<T, FAST>(hyd, zb_in, G_in, m0, sd, long(p.r0 - 1) * nx + p.cc);
        if (ed_active(up, cur)) fn = ex_edge_d<1, T, FAST>(up, cur);
    }
    for (int r = p.r0; r < p.r1; ++r) {
        const bool hasS = r + 1 < ny;
        ED<T> nxt = cur;
        if (hasS) nxt = ed_load<T, FAST>(hyd, zb_in, G_in, m0, sd, long(r + 1) * nx + p.cc);
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
            const long i = long(r) * nx + c;
            const T inc = FAST ? dtdx * div : dt * div / dx;
            T zb1 = cur.zb - inc, zl1 = zero;
            if constexpr (sizeof(T) == 4) {
                if (zb_lo_in) cb_sub(cur.zb, zb_lo_in[i], inc, zb1, zl1);
            }
            if (MODE == 0) {
                out.z_b[i] = zb1;
                if constexpr (sizeof(T) == 4) {
                    if (out.z_b_lo) out.z_b_lo[i] = zl1;
                }
                if (write_G) out.G[i] = cur.G;
            } else {
                const T hf = T(0.5);
                out.h[i] = hf * old.h[i] + hf * cur.h;
                out.hu[i] = hf * old.hu[i] + hf * cur.hu;
                out.hv[i] = hf * old.hv[i] + hf * cur.hv;
                out.G[i] = hf * old.G[i] + hf * cur.G;
                T zc = hf * old.z_b[i] + hf * zb1;
                if constexpr (sizeof(T) == 4) {
                    if (out.z_b_lo) {
                        T zcl;
                        cb_half_sum(old.z_b[i], old.z_b_lo[i], zb1, zl1, zc, zcl);
                        out.z_b_lo[i] = zcl;
                    }
                }
                out.z_b[i] = zc;
            }
        }
        fn = fy;
        cur = nxt;
    }
}

// ---------------------------------------------------------------------------
// CFL dt partial minima (roe_exner.py:257-310), one partial per block
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(SW *SWPB) void k_rx_dt_strip(F<T> s, const uint8_t *__restrict__ m0, int ny, int nx,
                                                          T dx, T *__restrict__ partial) {
    __shared__ T red[SWPB];
    const StripPos p = strip_pos(ny, nx);
    const T T_ = C<T>(DRY_TOL);
    T v = inf_v<T>();
    if (p.r0 < ny) {
        const long i0 = long(p.r0) * nx + p.cc;
        T h = s.h[i0], hu = s.hu[i0], hv = s.hv[i0], G = s.G[i0];
        int nod = m0[i0];
        for (int r = p.r0; r < p.r1; ++r) {
            const bool hasS = r + 1 < ny;
            T hS = h, huS = hu, hvS = hv, GS = G;
            int nodS = nod;
            if (hasS) {
                const long j = long(r + 1) * nx + p.cc;
                hS = s.h[j]; huS = s.hu[j]; hvS = s.hv[j]; GS = s.G[j]; nodS = m0[j];
            }
            const T hE = shdn(h), huE = shdn(hu), hvE = shdn(hv), GE = shdn(G);
            const int nodE = shdn(nod);
            // lanes 1..62 own the x edge (c | c+1) and the y edge (r | r+1) of their column
            if (p.out && p.c + 1 < nx && ((h > T_) || (hE > T_)) && !(nod || nodE))
                v = pmin(v, dt_edge<0>(h, hu, hv, G, hE, huE, hvE, GE, dx));
            if (p.out && hasS && ((h > T_) || (hS > T_)) && !(nod || nodS))
                v = pmin(v, dt_edge<1>(h, hu, hv, G, hS, huS, hvS, GS, dx));
            h = hS; hu = huS; hv = hvS; G = GS; nod = nodS;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    if (p.lane == 0) red[threadIdx.x / SW] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < SWPB; ++k) v = pmin(v, red[k]);
        partial[blockIdx.y * gridDim.x + blockIdx.x] = v;
    }
}

}  // namespace pxh
