// Plain Roe edge flux (solvers/kernels/roe.py:32-249) and cell CFL
// (roe.py:13-30), operation order as in the reference, quirks included:
// lambda_E from the already replaced lambda (roe.py:82-92), (dx/2)*lambda in
// the source fix (roe.py:180-181), beta3min from hi - alpha_1 (roe.py:173),
// live dry masks (roe.py:196-197), no active-edge masking.
#pragma once
#include "common.hpp"
#include "flux_roe_exner.hpp"

namespace pxh {

template <typename T>
struct RoeCell {
    T h, hu, hv, z, n;
};

template <int AX, typename T>
__device__ __forceinline__ void roe_edge(const RoeCell<T> &I, const RoeCell<T> &J, T dx, T P[3], T M[3]) {
    const T T_ = C<T>(DRY_TOL), g = C<T>(GRAV), V = C<T>(VEL_TOL), zero = T(0), one = T(1);
    T hi = I.h > T_ ? I.h : zero;
    T hj = J.h > T_ ? J.h : zero;
    T si = d_sqrt(hi), sj = d_sqrt(hj);
    T ui = hi > zero ? I.hu / hi : zero;
    T vi = hi > zero ? I.hv / hi : zero;
    T uj = hj > zero ? J.hu / hj : zero;
    T vj = hj > zero ? J.hv / hj : zero;
    T uhi, vhi, uhj, vhj;
    rot<AX>(ui, vi, uhi, vhi);
    rot<AX>(uj, vj, uhj, vhj);
    T hav = T(0.5) * (hi + hj);
    T ht = hav > T_ ? hav : zero;
    T ct = d_sqrt(g * ht);
    T ut = (uhi * si + uhj * sj) / (si + sj);
    T vt = (vhi * si + vhj * sj) / (si + sj);
    T l1 = ut - ct, l2 = ut, l3 = ut + ct;
    T ei = uhi - d_sqrt(g * hi), ej = uhj - d_sqrt(g * hj);
    T lE1 = zero;
    if ((ei < zero) && (ej > zero)) {
        l1 = ei * (ej - l1) / ((ej - ei) + V);
        lE1 = l1 - ei * (ej - l1) / ((ej - ei) + V);
    }
    ei = uhi + d_sqrt(g * hi);
    ej = uhj + d_sqrt(g * hj);
    T lE3 = zero;
    if ((ei < zero) && (ej > zero)) {
        l3 = ej * (l3 - ei) / ((ej - ei) + V);
        lE3 = l3 - ej * (l3 - ei) / ((ej - ei) + V);
    }
    // P_inv rows (roe.py:107-113)
    T den = l1 - l3 + V;
    T q00 = -l3 / den, q01 = one / den;
    T q10 = -vt / ct, q12 = one / ct;
    T q20 = l1 / den, q21 = -one / den;
    T dh = hj - hi;
    T dhu = hj * uhj - hi * uhi;
    T dhv = hj * vhj - hi * vhi;
    // einsum rows with their zero entries kept (0*x) so NaN/inf propagate as coded
    T a0 = q00 * dh + q01 * dhu + zero * dhv;
    T a1 = q10 * dh + zero * dhu + q12 * dhv;
    T a2 = q20 * dh + q21 * dhu + zero * dhv;
    // sources (roe.py:132-157)
    T dz = J.z - I.z;
    T di = hi + I.z, dj = hj + J.z;
    T dd = dj - di;
    T thra = (-g * ht) * dz;
    bool m1 = (dz >= zero) && (di < J.z);
    bool m2 = (dz < zero) && (dj < I.z);
    T dzt = m1 ? hi : (m2 ? hj : dz);
    T hr = (dz >= zero) ? hi : hj;
    T thrb = (-g * (hr - T(0.5) * d_abs(dzt))) * dzt;
    bool mt = (dz * dd >= zero) && (ut * dz > zero);
    T thrust = mt ? pmax(thra, thrb) : thrb;
    T nt = T(0.5) * (I.n + J.n);
    T sf = (sq(nt) * d_sqrt(sq(ut) + sq(vt)) * ut) / pmax(T_, d_pow(ht, C<T>(4.0 / 3.0)));
    T tau = g * ht * sf * dx;
    T tn = thrust - tau;
    T b0 = q00 * zero + q01 * tn + zero * zero;
    T b1 = q10 * zero + zero * tn + q12 * zero;
    T b2 = q20 * zero + q21 * tn + zero * zero;
    // beta fix (roe.py:169-189)
    T h1s = hi + a0 - (b0 / l1);
    T h3s = hj - a2 + (b2 / l3);
    T b1min = -(hi + a0) * d_abs(l1);
    T b3min = -(hi - a0) * d_abs(l3);
    T lmax = pmax(pmax(d_abs(l1), d_abs(l2)), d_abs(l3));
    T dtl = dx / lmax;
    T dt1s = (dx / T(2) * l1) * (hi / (hi - h1s + V));
    T dt3s = (dx / T(2) * l3) * (hj / (hj - h3s + V));
    if ((h1s < zero) && (h3s > zero) && (dt1s < dtl)) {
        b0 = (-b1min >= b3min) ? b1min : b0;
        b2 = -b0;
    }
    if ((h1s > zero) && (h3s < zero) && (dt3s < dtl)) {
        b2 = (-b3min >= b3min) ? b1min : b2;
        b0 = -b2;
    }
    bool mk1 = (hi < T_) && (h1s < zero);
    bool mk2 = (hj < T_) && (h3s < zero);
    T p0 = zero, p1 = zero, p2 = zero, n0 = zero, n1 = zero, n2 = zero;
    const T lam[3] = {l1, l2, l3};
    const T alp[3] = {a0, a1, a2};
    const T bet[3] = {b0, b1, b2};
    const T r0[3] = {one, zero, one};
    const T r1[3] = {l1, zero, l3};
    const T r2[3] = {vt, ct, vt};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T f = lam[k] * alp[k] - bet[k];
        T f0 = f * r0[k], f1 = f * r1[k], f2 = f * r2[k];
        bool pos = mk2 ? false : (mk1 ? true : (lam[k] > zero));
        bool neg = mk2 ? true : (mk1 ? false : (lam[k] <= zero));
        if (pos) { p0 += f0; p1 += f1; p2 += f2; }
        if (neg) { n0 += f0; n1 += f1; n2 += f2; }
    }
    const T lE[3] = {lE1, zero, lE3};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T f = lE[k] * alp[k];
        T f0 = f * r0[k], f1 = f * r1[k], f2 = f * r2[k];
        if (lE[k] > zero) { p0 += f0; p1 += f1; p2 += f2; }
        if (lE[k] <= zero) { n0 += f0; n1 += f1; n2 += f2; }
    }
    unrot<AX>(p1, p2);
    unrot<AX>(n1, n2);
    P[0] = p0; P[1] = p1; P[2] = p2;
    M[0] = n0; M[1] = n1; M[2] = n2;
}

}  // namespace pxh
