// One-cell pieces of the per-kernel Roe-Exner stage shared by the forward
// kernels (pyexner_hip.hip) and their adjoints (adjoint.hip): the hydro update
// of a cell, the momentum-correction masks, the Exner edge activity test, the
// momentum transform of a ring-fill entry, and the host-side argument checks.
#pragma once
#include "flux_roe_exner.hpp"

namespace pxh {

template <typename T>
static F<T> cast_fields(const pxh_fields *f) {
    F<T> r;
    r.h = static_cast<T *>(f->h);
    r.hu = static_cast<T *>(f->hu);
    r.hv = static_cast<T *>(f->hv);
    r.z = static_cast<T *>(f->z);
    r.z_b = static_cast<T *>(f->z_b);
    r.n = static_cast<T *>(f->n);
    r.n_b = static_cast<T *>(f->n_b);
    r.G = static_cast<T *>(f->G);
    r.z_b_lo = sizeof(T) == 4 ? static_cast<T *>(f->z_b_lo) : nullptr;
    return r;
}

template <typename T>
__device__ __forceinline__ HydroCell<T> load_hc(const F<T> &s, int i) {
    HydroCell<T> c;
    c.h = s.h[i]; c.hu = s.hu[i]; c.hv = s.hv[i];
    c.zb = s.z_b[i]; c.z = s.z[i]; c.n = s.n_b[i];
    return c;
}

// roe_solve_2D (roe_exner.py:535-612) for cell (r, c): the flux sum
// F = M_x[r,c] + P_x[r,c-1] + M_y[r,c] + P_y[r-1,c] in that order over the
// active edges, then q - dt F / dx.
template <typename T>
__device__ __forceinline__ void rx_hydro_cell(const F<T> &in, const uint8_t *__restrict__ m0, T dt, int ny, int nx,
                                              T dx, int r, int c, T &h1, T &hu1, T &hv1) {
    const int i = r * nx + c;
    const T T_ = C<T>(DRY_TOL);
    const HydroCell<T> me = load_hc(in, i);
    const bool mme = m0[i] != 0;
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    T P[3], M[3];
    if (c + 1 < nx) {  // east edge: this cell is side i -> upwM
        const int j = i + 1;
        const HydroCell<T> o = load_hc(in, j);
        if (((me.h > T_) || (o.h > T_)) && !(mme || m0[j])) {
            roe_exner_edge<0>(me, o, dx, P, M);
            acc0 = acc0 + M[0]; acc1 = acc1 + M[1]; acc2 = acc2 + M[2];
        }
    }
    if (c > 0) {  // west edge: this cell is side j -> upwP
        const int j = i - 1;
        const HydroCell<T> o = load_hc(in, j);
        if (((o.h > T_) || (me.h > T_)) && !(m0[j] || mme)) {
            roe_exner_edge<0>(o, me, dx, P, M);
            acc0 = acc0 + P[0]; acc1 = acc1 + P[1]; acc2 = acc2 + P[2];
        }
    }
    if (r + 1 < ny) {  // south edge (row r+1): this cell is side i -> upwM
        const int j = i + nx;
        const HydroCell<T> o = load_hc(in, j);
        if (((me.h > T_) || (o.h > T_)) && !(mme || m0[j])) {
            roe_exner_edge<1>(me, o, dx, P, M);
            acc0 = acc0 + M[0]; acc1 = acc1 + M[1]; acc2 = acc2 + M[2];
        }
    }
    if (r > 0) {  // north edge: this cell is side j -> upwP
        const int j = i - nx;
        const HydroCell<T> o = load_hc(in, j);
        if (((o.h > T_) || (me.h > T_)) && !(m0[j] || mme)) {
            roe_exner_edge<1>(o, me, dx, P, M);
            acc0 = acc0 + P[0]; acc1 = acc1 + P[1]; acc2 = acc2 + P[2];
        }
    }
    h1 = me.h - dt * acc0 / dx;
    hu1 = me.hu - dt * acc1 / dx;
    hv1 = me.hv - dt * acc2 / dx;
}

// momentum_corrections (roe_exner.py:14-62) at interior cell i: whether hu
// and hv survive.  h1 is the hydro update before the ring fill; z + zb the
// bed before this step's Exner update.
template <typename T>
__device__ __forceinline__ void momcorr_keep(const T *__restrict__ h1, const T *__restrict__ z,
                                             const T *__restrict__ zb, const uint8_t *__restrict__ m0, int nx, int i,
                                             bool &keep_u, bool &keep_v) {
    const T T_ = C<T>(DRY_TOL);
    const int iW = i - 1, iE = i + 1, iN = i - nx, iS = i + nx;
    const T hc = h1[i];
    const T eta = hc + (z[i] + zb[i]);
    const T zW = z[iW] + zb[iW], zE = z[iE] + zb[iE];
    const T zN = z[iN] + zb[iN], zS = z[iS] + zb[iS];
    const bool dryW = h1[iW] < T_, dryE = h1[iE] < T_;
    const bool dryN = h1[iN] < T_, dryS = h1[iS] < T_;
    const bool stop_u = ((eta < zE) && dryE) || ((eta < zW) && dryW) || m0[iE] || m0[iW];
    const bool stop_v = ((eta < zN) && dryN) || ((eta < zS) && dryS) || m0[iN] || m0[iS];
    const bool active = (hc > T_) && !m0[i];
    keep_u = active && !stop_u;
    keep_v = active && !stop_v;
}

template <typename T>
__device__ __forceinline__ ExnerCell<T> load_ec(const F<T> &s, const T *zb, int i) {
    ExnerCell<T> c;
    c.h = s.h[i]; c.hu = s.hu[i]; c.hv = s.hv[i]; c.zb = zb[i]; c.G = s.G[i];
    return c;
}

// The active-edge rule of exner_solve_2D (roe_exner.py:681-690).
template <typename T>
__device__ __forceinline__ bool exner_active(const ExnerCell<T> &a, const ExnerCell<T> &b, bool ma, bool mb) {
    const T T_ = C<T>(DRY_TOL), S_ = C<T>(SED_TOL);
    const bool stop = (ma || mb) || ((a.zb < S_) && (b.zb < S_));
    const bool depth = (a.h > T_) || (b.h > T_);
    return depth && !stop;
}

// The momentum transform of one handler of a ring-fill entry
// (roeexner_reflective.py:31-59, roe_reflective.py:34-39).  Both reflections
// are symmetric linear maps of (hu, hv).
template <typename T>
__device__ __forceinline__ void bc_transform(const pxh_bc_handler &hd, T &hu, T &hv) {
    const T nx = static_cast<T>(hd.nx), ny = static_cast<T>(hd.ny);
    if (hd.kind == PXH_BC_REFLECT_RX) {
        const T mn = hu * nx + hv * ny;
        const T mt = -hu * ny + hv * nx;
        const T mr = -mn;
        const T u2 = mr * nx - mt * ny;
        const T v2 = mr * ny + mt * nx;
        hu = u2;
        hv = v2;
    } else if (hd.kind == PXH_BC_REFLECT_ROE) {
        hu = hu * (T(1) - T(2) * nx * nx);
        hv = hv * (T(1) - T(2) * ny * ny);
    }
}

// The field of a ring-fill group (PXH_G_HUV: hu; its hv is q = 1).
template <typename T>
__device__ __forceinline__ T *group_field(const F<T> &s, int group, int q = 0) {
    switch (group) {
        case PXH_G_H: return s.h;
        case PXH_G_HUV: return q ? s.hv : s.hu;
        case PXH_G_Z: return s.z;
        case PXH_G_ZB: return s.z_b;
        case PXH_G_N: return s.n;
        case PXH_G_NB: return s.n_b;
        case PXH_G_G: return s.G;
        default: return nullptr;
    }
}

// The depth a ConstantOutflux boundary sets (roeexner_constantoutflux.py:24-56):
// five Newton steps on q/h + 2 sqrt(g h) = un_i + 2 sqrt(g h_i), the outgoing
// Riemann invariant of the interior cell, from h = h_i.
template <typename T>
__device__ __forceinline__ T outflux_depth(T h_src, T hu_src, T hv_src, T q, T nx, T ny) {
    const T g = C<T>(GRAV), fl = C<T>(1e-8);
    const T h_i = pmax(h_src, fl);
    const T u_i = hu_src / h_i, v_i = hv_src / h_i;
    const T un_i = u_i * nx + v_i * ny;
    const T c_i = d_sqrt(g * h_i);
    const T Rp = un_i + T(2) * c_i;
    T h = h_i;
    for (int k = 0; k < 5; ++k) {
        h = pmax(h, fl);
        const T c = d_sqrt(g * h);
        const T f = q / h + T(2) * c - Rp;
        const T df = -q / (h * h) + g / pmax(c, fl);
        h = pmax(h - f / df, fl);
    }
    return h;
}

// A value entry of the ring fill (pxh_bc_value); the compensated-bed low part
// of a set bed is zero, of an offset bed the source's.
template <typename T>
// sq: the element stride of s.h, hu, hv, z_b (4 for the packed records of
// PXH_PACKED_STAGE, else 1); the other fields are plain arrays.
__device__ __forceinline__ void bc_value_apply(const F<T> &s, const pxh_bc_entry &e, const pxh_bc &bc, int sq = 1) {
    if (e.chain[0] >= bc.n_values) return;  // no such value: the entry is ignored
    const pxh_bc_value v = bc.values[e.chain[0]];
    const int t = e.target, src = e.src;
    T *f = group_field(s, e.group);
    if (!f) return;
    const bool q4 = e.group == PXH_G_H || e.group == PXH_G_HUV || e.group == PXH_G_ZB;
    const long ft = q4 ? long(t) * sq : t, fs = q4 ? long(src) * sq : src;  // f's elements of t and src
    const long tq = long(t) * sq, sqi = long(src) * sq;                      // h, hu, hv's
    switch (v.kind) {
        case PXH_BV_CONST:
            if (e.group == PXH_G_HUV) {
                if (v.a == v.a) s.hu[tq] = static_cast<T>(v.a);
                if (v.b == v.b) s.hv[tq] = static_cast<T>(v.b);
            } else {
                f[ft] = static_cast<T>(v.a);
                if (e.group == PXH_G_ZB && s.z_b_lo) s.z_b_lo[t] = T(0);
            }
            break;
        case PXH_BV_AFFINE_T:
            f[ft] = static_cast<T>(v.b) * static_cast<T>(bc.time) + static_cast<T>(v.a);
            if (e.group == PXH_G_ZB && s.z_b_lo) s.z_b_lo[t] = T(0);
            break;
        case PXH_BV_OFFSET:
            f[ft] = f[fs] + static_cast<T>(v.a);
            if (e.group == PXH_G_ZB && s.z_b_lo) s.z_b_lo[t] = s.z_b_lo[src];
            break;
        case PXH_BV_OUTFLUX:
            if (e.group == PXH_G_H)
                s.h[tq] = outflux_depth<T>(s.h[sqi], s.hu[sqi], s.hv[sqi], static_cast<T>(v.a),
                                           static_cast<T>(v.b), static_cast<T>(v.c));
            break;
        default: break;
    }
}

// ---------------------------------------------------------------------------
// host-side argument checks
// ---------------------------------------------------------------------------
static inline bool grid_ok(const pxh_grid *g) {
    if (!g) return false;
    const int rows = PXH_STRIP_ROWS(g->flags);
    return g->ny >= 2 && g->nx >= 2 && (g->dtype == PXH_F32 || g->dtype == PXH_F64) &&
           (long(g->ny) * g->nx) < (1L << 31) && (rows == 0 || (rows >= 8 && rows <= 128));
}

// A boundary program (or none) the ring-fill launch can run: with
// PXH_BC_GROUPED, group_start must be non-decreasing from 0 to n_entries.
static inline bool bc_ok(const pxh_bc *bc) {
    if (!bc) return true;
    if (bc->n_handlers < 0 || bc->n_handlers > PXH_MAX_HANDLERS || bc->n_entries < 0) return false;
    if (bc->n_values < 0 || bc->n_values > PXH_MAX_VALUES) return false;
    if (bc->n_entries > 0 && !bc->entries) return false;
    if (bc->n_levels < 0 || bc->n_levels > PXH_MAX_LEVELS) return false;
    if (bc->n_levels > 1) {
        if (bc->level_start[0] != 0 || bc->level_start[bc->n_levels] != bc->n_entries) return false;
        for (int q = 0; q < bc->n_levels; ++q)
            if (bc->level_start[q + 1] < bc->level_start[q]) return false;
    } else if (bc->flags & PXH_BC_GROUPED) {
        if (bc->group_start[0] != 0 || bc->group_start[PXH_N_GROUPS] != bc->n_entries) return false;
        for (int q = 0; q < PXH_N_GROUPS; ++q)
            if (bc->group_start[q + 1] < bc->group_start[q]) return false;
    }
    return true;
}

// records: accept a view of four-element records (hu, hv, z_b one, two and
// three elements past h), which the stage and frame sweeps and the ring fill
// read; elsewhere an hu that starts within one element past h overlaps h.
static inline bool fields_ok(const pxh_fields *f, bool exner, bool records = false) {
    if (!f || !f->h || !f->hu || !f->hv || !f->z || !f->n) return false;
    const long d = static_cast<const char *>(f->hu) - static_cast<const char *>(f->h);
    if (!records && d > 0 && d <= 8) return false;
    if (exner && (!f->z_b || !f->n_b || !f->G)) return false;
    return true;
}

}  // namespace pxh
