// Roe-Exner edge fluxes, the Exner bedload flux, MPM transport coefficient and
// the coupled-cubic CFL speed, written for one edge (or cell) per call.
//
// Operation order follows solvers/kernels/roe_exner.py exactly so that the
// double build reproduces the float64 restatement to round-off: no operand is
// re-associated and the library is built with -ffp-contract=off.
#pragma once
#include "common.hpp"

namespace pxh {

template <typename T>
struct HydroCell {
    T h, hu, hv, zb, z, n;
};

// Rotation into the edge frame.  AX = 0: x edge, normal (1, 0); AX = 1: y
// edge, normal (0, -1) (roe_exner.py:576-577).  u*1 + v*0 == u and
// u*0 + v*(-1) == -v exactly, so the general formula reduces to these.
template <int AX, typename T>
__device__ __forceinline__ void rot(T u, T v, T &un, T &ut) {
    if (AX == 0) { un = u; ut = v; }
    else { un = -v; ut = u; }
}
// Tk_inv rotate-back (roe_exner.py:526-531): x: (f0, f1, f2); y: (f0, f2, -f1).
template <int AX, typename T>
__device__ __forceinline__ void unrot(T &f1, T &f2) {
    if (AX == 1) { T t = f1; f1 = f2; f2 = -t; }
}

// roe_solver (roe_exner.py:312-533).  Returns upwP (to cell j) and upwM (to
// cell i) in the global frame.  dt-independent.
template <int AX, typename T>
__device__ __forceinline__ void roe_exner_edge(const HydroCell<T> &I, const HydroCell<T> &J, T dx,
                                               T P[3], T M[3]) {
    const T T_ = C<T>(DRY_TOL), g = C<T>(GRAV), V = C<T>(VEL_TOL), zero = T(0);
    T hi = pmax(I.h, T_), hj = pmax(J.h, T_);
    T ui = hi > T_ ? I.hu / hi : zero;
    T vi = hi > T_ ? I.hv / hi : zero;
    T uj = hj > T_ ? J.hu / hj : zero;
    T vj = hj > T_ ? J.hv / hj : zero;
    T uhi, vhi, uhj, vhj;
    rot<AX>(ui, vi, uhi, vhi);
    rot<AX>(uj, vj, uhj, vhj);
    T si = d_sqrt(hi), sj = d_sqrt(hj);
    T den = si + sj;
    T ht = T(0.5) * (hi + hj);
    T ct = d_sqrt(g * ht);
    T ut = (uhi * si + uhj * sj) / den;
    T vt = (vhi * si + vhj * sj) / den;
    T l1r = ut - ct, l2 = ut, l3r = ut + ct;
    // Harten-Hyman (roe_exner.py:357-373)
    T ei = uhi - d_sqrt(g * hi), ej = uhj - d_sqrt(g * hj);
    T lE1 = zero, l1 = l1r;
    if ((ei < zero) && (ej > zero)) {
        lE1 = l1r - ei * (ej - l1r) / (ej - ei);
        l1 = ei * (ej - l1r) / (ej - ei);
    }
    ei = uhi + d_sqrt(g * hi);
    ej = uhj + d_sqrt(g * hj);
    T lE3 = zero, l3 = l3r;
    if ((ei < zero) && (ej > zero)) {
        lE3 = l3r - ej * (l3r - ei) / (ej - ei);
        l3 = ej * (l3r - ei) / ((ej - ei) + V);
    }
    // wave strengths (roe_exner.py:399-412)
    T dh = hj - hi;
    T dhu = hj * uhj - hi * uhi;
    T dhv = hj * vhj - hi * vhi;
    T dn2 = T(-2) * ct;
    T a1 = (-l3r * dh + dhu) / dn2;
    T a2 = (-vt * dh + dhv) / ct;
    T a3 = (l1r * dh - dhu) / dn2;
    // bed thrust (roe_exner.py:421-432)
    const T S_ = C<T>(SED_TOL), SF = C<T>(SED_FLOOR);
    T zbi = I.zb > S_ ? I.zb : SF;
    T zbj = J.zb > S_ ? J.zb : SF;
    T dz = (J.z + zbj) - (I.z + zbi);
    T thrust = (-g * ht) * dz;
    // Manning friction (roe_exner.py:455-457); ntilde == 0 gives tau == 0 exactly
    T nt = T(0.5) * (I.n + J.n);
    T tau = zero;
    if (nt != zero) {
        T sf = (sq(nt) * d_sqrt(sq(ut) + sq(vt)) * ut) / pmax(T_, d_pow(ht, C<T>(4.0 / 3.0)));
        tau = g * ht * sf * dx;
    }
    T b1 = (thrust - tau) / dn2;
    // upwinding (roe_exner.py:484-524); the h*-star dry masks are dead (h >= DRY_TOL)
    // each wave goes to upwP when lambda > 0 and to upwM when lambda <= 0
    // (a NaN lambda contributes to neither, as jnp.where does)
    T p0 = zero, p1 = zero, p2 = zero, m0 = zero, m1 = zero, m2 = zero;
    auto acc = [&](T lam, T f0, T f1, T f2) {
        if (lam > zero) { p0 += f0; p1 += f1; p2 += f2; }
        if (lam <= zero) { m0 += f0; m1 += f1; m2 += f2; }
    };
    {  // wave 1: r = (1, l1, vt)
        T f = l1 * a1 - b1;
        acc(l1, f * T(1), f * l1, f * vt);
    }
    {  // wave 2: r = (0, 0, ct), beta = 0
        T f = l2 * a2 - zero;
        acc(l2, f * zero, f * zero, f * ct);
    }
    {  // wave 3: r = (1, l3, vt), beta = -b1
        T f = l3 * a3 - (-b1);
        acc(l3, f * T(1), f * l3, f * vt);
    }
    {  // entropy wave 1
        T f = lE1 * a1;
        acc(lE1, f * T(1), f * l1, f * vt);
    }
    {  // entropy wave 2: lambda_E2 == 0
        T f = zero * a2;
        acc(zero, f * zero, f * zero, f * ct);
    }
    {  // entropy wave 3
        T f = lE3 * a3;
        acc(lE3, f * T(1), f * l3, f * vt);
    }
    unrot<AX>(p1, p2);
    unrot<AX>(m1, m2);
    P[0] = p0; P[1] = p1; P[2] = p2;
    M[0] = m0; M[1] = m1; M[2] = m2;
}

template <typename T>
struct ExnerCell {
    T h, hu, hv, zb, G;
};

// exner_solver (roe_exner.py:614-663): ACM upwind bedload flux on one edge.
template <int AX, typename T>
__device__ __forceinline__ T exner_edge(const ExnerCell<T> &I, const ExnerCell<T> &J) {
    const T T_ = C<T>(DRY_TOL), S_ = C<T>(SED_TOL), zero = T(0);
    T hi = I.h > T_ ? I.h : C<T>(DRY_FLOOR);
    T hj = J.h > T_ ? J.h : C<T>(DRY_FLOOR);
    T zbi = I.zb > S_ ? I.zb : C<T>(SED_FLOOR);
    T zbj = J.zb > S_ ? J.zb : C<T>(SED_FLOOR);
    T si = d_sqrt(hi), sj = d_sqrt(hj);
    T ui = hi > T_ ? I.hu / hi : zero;
    T vi = hi > T_ ? I.hv / hi : zero;
    T uj = hj > T_ ? J.hu / hj : zero;
    T vj = hj > T_ ? J.hv / hj : zero;
    T uhi = (AX == 0) ? ui : -vi;
    T uhj = (AX == 0) ? uj : -vj;
    T ut = (uhi * si + uhj * sj) / (si + sj);
    T umi = sq(ui) + sq(vi);
    T umj = sq(uj) + sq(vj);
    T qbi = I.G * umi * uhi;
    T qbj = J.G * umj * uhj;
    T gt = T(0.5) * (I.G + J.G);
    T dz = zbj - zbi;
    T dq = gt * umj * uhj - gt * umi * uhi;
    T l4 = (d_abs(dz) > S_) ? dq / dz : ut;
    T ci = (gt - I.G) * umi * uhi;
    T cj = (gt - J.G) * umj * uhj;
    T qni = qbi + ci;
    T qnj = qbj + cj;
    return (l4 >= zero) ? qni : qnj;
}

// compute_G (roe_exner.py:85-130) for one cell, MPM mode.  nodata is mask[0].
template <typename T>
__device__ __forceinline__ T mpm_G(T h, T hu, T hv, bool nodata, const pxh_sediments &sd) {
    const T T_ = C<T>(DRY_TOL), zero = T(0);
    h = pmax(h, T_);
    T u = hu / h, v = hv / h;
    T vel2 = sq(u) + sq(v);
    T cbh = d_cbrt(h);
    T sqh = d_sqrt(h);
    T G = zero;
    for (int k = 0; k < sd.n_sed; ++k) {
        T th = (C<T>(sd.A[k]) * vel2) / (C<T>(sd.B[k]) * cbh);
        th = pmax(th, zero);
        T dth = pmax(th - C<T>(sd.theta_c[k]), zero);
        T g1 = C<T>(sd.C[k]) / (C<T>(sd.D[k]) * sqh);
        T g2 = T(8) * d_sqrt(dth) / (d_pow(th, T(1.5)) + C<T>(1e-5));
        T c = C<T>(sd.frac[k]) * g1 * g2 * dth;
        G = (k == 0) ? c : G + c;
    }
    G = G / C<T>(sd.one_minus_por);
    if (nodata) G = zero;
    if (!(h > T_)) G = zero;
    return G;
}

// compute_G_factory constant branch (roe_exner.py:136-140).
template <typename T>
__device__ __forceinline__ T const_G(T h, T Gprev, bool nodata) {
    T G = nodata ? T(0) : Gprev;
    return (h > C<T>(DRY_TOL)) ? G : T(0);
}

// Largest and smallest of the three values cos((acos(x) + 2 pi k) / 3), with
// acos's argument clipped to [-1, 1] as _get_lambda does (roe_exner.py:185).
// They are the extreme roots of 4 t^3 - 3 t - x = 0.  For x >= 0 the largest
// root lies in [sqrt(3)/2, 1], where the derivative 12 t^2 - 3 >= 6, so three
// Newton steps from a quadratic fit (error < 2e-3) reach double precision; the
// smallest root follows from dividing out that factor:
// t' = (-t - sqrt(3 (1 - t)(1 + t))) / 2, accurate also next to the double
// root at x -> 1.  Negative x uses the mirror t(-x) = -t(x).
__host__ __device__ __forceinline__ void trisect(double x, double &tmin, double &tmax) {
    if (x >= 1.0) { tmax = 1.0; tmin = -0.5; return; }
    if (x <= -1.0) { tmax = 0.5; tmin = -1.0; return; }
    const bool neg = x < 0.0;
    const double xa = neg ? -x : x;
    double t = 0.8660254037844386 + xa * (0.16666666666666666 - 0.03269207045111198 * xa);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double t2 = t * t;
        t = t - ((4.0 * t2 - 3.0) * t - xa) / (12.0 * t2 - 3.0);
    }
    const double other = 0.5 * (-t - __builtin_sqrt(3.0 * (1.0 - t) * (1.0 + t)));
    if (neg) { tmin = -t; tmax = -other; }
    else { tmax = t; tmin = other; }
}

// _get_lambda (roe_exner.py:163-209), evaluated in double: the coded cubic's
// depressed form, lambda = 2 sqrt(-p/3) t + shift with t from trisect.
__device__ __forceinline__ void get_lambda(double ut, double at, double ct, double &lmin, double &lmax) {
    if (at < 1e-5) {
        lmin = ut - ct;
        lmax = ut + ct;
        return;
    }
    const double a = -1.0;
    double b = -2.0 * ut;
    double c = (1.0 + at) * (ct * ct) - ut * ut;
    double d = ((-(ct * ct)) * at) * ut;
    double p = ((3.0 * a) * c - b * b) / 3.0;
    double q = ((2.0 * cube(b) - ((9.0 * a) * b) * c) + 27.0 * d) / -27.0;
    double x = (3.0 * q) / (2.0 * p) * __builtin_sqrt(-3.0 / p);
    double tmin, tmax;
    trisect(x, tmin, tmax);
    double st = 2.0 * __builtin_sqrt(-p / 3.0);
    double sh = -b / (3.0 * a);
    lmax = st * tmax + sh;
    lmin = st * tmin + sh;
}

// compute_dt per edge (roe_exner.py:211-255).
template <int AX, typename T>
__device__ __forceinline__ T dt_edge(T Ih, T Ihu, T Ihv, T IG, T Jh, T Jhu, T Jhv, T JG, T dx) {
    const T T_ = C<T>(DRY_TOL), g = C<T>(GRAV);
    T hi = pmax(Ih, T_), hj = pmax(Jh, T_);
    T si = d_sqrt(hi), sj = d_sqrt(hj);
    T den = si + sj;
    T ui = Ihu / hi, vi = Ihv / hi, uj = Jhu / hj, vj = Jhv / hj;
    T uhi, vhi, uhj, vhj;
    rot<AX>(ui, vi, uhi, vhi);
    rot<AX>(uj, vj, uhj, vhj);
    T ht = T(0.5) * (hi + hj);
    T ut = (uhi * si + uhj * sj) / den;
    T ct = d_sqrt(g * ht);
    T gt = T(0.5) * (IG + JG);
    T at = gt * (sq(uhi) + uhi * uhj + sq(uhj) + vhi * vhj) / (si * sj);
    T l1 = ut - ct, l3 = ut + ct;
    double lmin, lmax;
    get_lambda(static_cast<double>(ut), static_cast<double>(at), static_cast<double>(ct), lmin, lmax);
    T rs = pmax(d_abs(static_cast<T>(lmin)), d_abs(static_cast<T>(lmax)));
    T ss = pmax(d_abs(l1), d_abs(l3));
    return (rs < C<T>(VEL_TOL)) ? dx / ss : dx / rs;
}

}  // namespace pxh
