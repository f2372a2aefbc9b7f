// Adjoint (reverse-mode) kernels of one Roe-Exner stage, of the CFL dt and of
// the initial compute_G / compute_n, in double.  They transpose the per-kernel
// forward of one stage (roe_exner_solver.py:80-116):
//
//   h1 = hydro(u)                      roe_solve_2D          rx_hydro_cell
//   (hu2, hv2) = keep * (hu1, hv1)     momentum_corrections  momcorr_keep
//   fill (h, hu, hv)                   BoundaryManager.apply
//   G = G(h2, hu2, hv2)                compute_G             mpm_G
//   zb' = zb - dt div F(h2, G, zb)/dx  exner_solve_2D        exner_edge
//   fill (z_b, G)                      BoundaryManager.apply
//
// in reverse.  Each cell thread gathers what its own inputs receive through
// the (up to four) edges touching it; an edge's derivatives with respect to
// that cell come from evaluating the forward edge function on Dual values
// seeded at the cell, so the branches are the ones the forward took.  Masks
// and branch selections (active edges, momentum-correction stops, upwind
// sides, clamps) are piecewise constant and carry no derivative.
//
// Sediment parameters (first sediment class): derivatives with respect to
// A = np^2, B = (s-1) d50, C = np^3 sqrt(g), 1 - porosity and the bed Manning
// coefficient nb = d50m^(1/6) / 21.1 (compute_n, roe_exner.py:64-73) are
// accumulated; the host chains them to d50 and porosity.  The erosion and
// deposition coefficients k_e, k_d never enter compute_G or the fluxes
// (roe_exner.py:85-130), so their derivatives are zero.
#pragma once
#include <type_traits>

#include "cell.hpp"
#include "dual.hpp"

namespace pxh {

using D4 = Dual<4>;
using D5 = Dual<5>;
using D8 = Dual<8>;
using AxX = std::integral_constant<int, 0>;
using AxY = std::integral_constant<int, 1>;

constexpr int ADJ_BX = 64, ADJ_BY = 4;
constexpr int ADJ_NPAR = 5;  // A, B, C, 1 - porosity, nb

// mpm_G (roe_exner.py:85-130) with the first class's A, B, C and
// 1 - porosity given as values of T: the same operations in the same order.
template <typename T>
__device__ __forceinline__ T mpm_G_par(T h, T hu, T hv, bool nodata, const pxh_sediments &sd, T A0, T B0, T C0,
                                       T omp) {
    const T T_ = C<T>(DRY_TOL), zero = T(0);
    h = pmax(h, T_);
    T u = hu / h, v = hv / h;
    T vel2 = sq(u) + sq(v);
    T cbh = d_cbrt(h);
    T sqh = d_sqrt(h);
    T G = zero;
    for (int k = 0; k < sd.n_sed; ++k) {
        const T A = k == 0 ? A0 : C<T>(sd.A[k]), B = k == 0 ? B0 : C<T>(sd.B[k]);
        const T Cc = k == 0 ? C0 : C<T>(sd.C[k]);
        T th = (A * vel2) / (B * cbh);
        th = pmax(th, zero);
        T dth = pmax(th - C<T>(sd.theta_c[k]), zero);
        T g1 = Cc / (C<T>(sd.D[k]) * sqh);
        T g2 = T(8) * d_sqrt(dth) / (d_pow(th, T(1.5)) + C<T>(1e-5));
        T c = C<T>(sd.frac[k]) * g1 * g2 * dth;
        G = (k == 0) ? c : G + c;
    }
    G = G / omp;
    if (nodata) G = zero;
    if (!(h > T_)) G = zero;
    return G;
}

// Sum of v over the block (ADJ_BX * ADJ_BY = 256 threads) into *slot.
__device__ __forceinline__ void block_sum_to(double v, double *slot) {
    __shared__ double red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    __syncthreads();
    if ((t & 63) == 0) red[t / 64] = v;
    __syncthreads();
    if (t == 0) *slot = red[0] + red[1] + red[2] + red[3];
}

// Ring fill transposed for the groups in `groups`: lam[src] += T^T lam[target],
// lam[target] = 0.  A program's targets are never its sources, so this is
// the transpose of the forward fill whatever the order.  The momentum
// transform of an entry is a product of symmetric reflections, so its
// transpose applies the same reflections in reverse order.  One pass: an
// entry moves its (transformed) target adjoint to its source and zeroes the
// target; no cell is both (a program of one dependency level), so no entry
// reads what another writes.  With src_sorted (entries of one group ordered
// by src) the first entry of each run of equal sources takes the run's
// targets and adds their sum, in entry order, so the result does not depend
// on scheduling; otherwise every entry adds its own with an atomic.
__device__ __forceinline__ double *adj_field(const F<double> &lam, int group, int q) {
    switch (group) {
        case PXH_G_H: return q ? nullptr : lam.h;
        case PXH_G_ZB: return q ? nullptr : lam.z_b;
        case PXH_G_NB: return q ? nullptr : lam.n_b;
        case PXH_G_G: return q ? nullptr : lam.G;
        case PXH_G_HUV: return q ? lam.hv : lam.hu;
        default: return nullptr;
    }
}

// Entry e's target adjoint, transformed for its source, into (a, b); the
// target zeroed.
__device__ __forceinline__ void bc_take(const F<double> &lam, const BcEnt &be, const pxh_bc &bc, double &a,
                                        double &b) {
    a = 0.0;
    b = 0.0;
    const pxh_bc_entry e = be.first();
    double *f0 = adj_field(lam, e.group, 0), *f1 = adj_field(lam, e.group, 1);
    const int t = e.target;
    if (e.nchain & PXH_BC_VALUE) {
        // a set value carries no derivative from the state (the target's
        // adjoint stops there); an offset passes it on like a copy; a NaN
        // momentum component leaves the cell as it was (identity)
        pxh_bc_value v{};
        if (e.chain[0] < bc.n_values) v = bc.values[e.chain[0]];
        const bool keep0 = v.kind == PXH_BV_CONST && e.group == PXH_G_HUV && !(v.a == v.a);
        const bool keep1 = v.kind == PXH_BV_CONST && e.group == PXH_G_HUV && !(v.b == v.b);
        if (f0 && !keep0) {
            if (v.kind == PXH_BV_OFFSET) a = f0[t];
            f0[t] = 0.0;
        }
        if (f1 && !keep1) f1[t] = 0.0;
        return;
    }
    if (f0) {
        a = f0[t];
        f0[t] = 0.0;
    }
    if (f1) {
        b = f1[t];
        f1[t] = 0.0;
    }
    if (e.group == PXH_G_HUV)
        for (int q = e.nchain - 1; q >= 0; --q) bc_transform<double>(bc.handlers[be.chain(q)], a, b);
}

__global__ __launch_bounds__(256) void k_adj_bc_move(F<double> lam, const pxh_bc_entry *__restrict__ ent, int n,
                                                     pxh_bc bc, uint32_t groups, int src_sorted) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const BcEnt be = bc_ent(ent, k);
    const unsigned grp = be.group();
    const int src = be.src();
    if (!((groups >> grp) & 1u)) return;
    double *f0 = adj_field(lam, int(grp), 0), *f1 = adj_field(lam, int(grp), 1);
    if (!f0 && !f1) return;
    if (src_sorted && k > 0) {
        const BcEnt prev = bc_ent(ent, k - 1);
        if (prev.group() == grp && prev.src() == src) return;  // not the run's first
    }
    double a, b;
    bc_take(lam, be, bc, a, b);
    if (src_sorted) {
        for (int j = k + 1; j < n; ++j) {
            const BcEnt nj = bc_ent(ent, j);
            if (nj.group() != grp || nj.src() != src) break;
            double aj, bj;
            bc_take(lam, nj, bc, aj, bj);
            a += aj;
            b += bj;
        }
    }
    if (f0) atomicAdd(&f0[src], a);
    if (f1) atomicAdd(&f1[src], b);
}

// The hydro update before the ring fill at the h targets of the program (the
// filled h equals it everywhere else): h1[target] = raw update.
__global__ __launch_bounds__(256) void k_adj_h1_targets(F<double> in, const uint8_t *__restrict__ m0,
                                                        const double *__restrict__ dtp, int ny, int nx, double dx,
                                                        const pxh_bc_entry *__restrict__ ent, int n,
                                                        double *__restrict__ h1) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const pxh_bc_entry e = ent[k];
    if (e.group != PXH_G_H) return;
    const int r = e.target / nx, c = e.target - r * nx;
    double a, b, d;
    rx_hydro_cell<double>(in, m0, *dtp, ny, nx, dx, r, c, a, b, d);
    h1[e.target] = a;
}

// ---------------------------------------------------------------------------
// Exner update and compute_G transposed.  hyd: the stage output's filled
// hydro (h2, hu2, hv2); zb_in: the stage input bed.  lam_out.z_b / lam_out.G:
// the adjoints of the Exner update and of compute_G before the final fill.
// Adds into lam_hyd (may alias lam_out's h, hu, hv), lam_zb_in (not aliasing
// lam_out.z_b), the per-block d/d dt partials and, with want_par, the
// per-block sediment-parameter partials (4 per block).
// ---------------------------------------------------------------------------
__device__ __forceinline__ ExnerCell<double> ec_at(const F<double> &hyd, const double *zb, const pxh_sediments &sd,
                                                  const uint8_t *m0, int i) {
    ExnerCell<double> c;
    c.h = hyd.h[i]; c.hu = hyd.hu[i]; c.hv = hyd.hv[i]; c.zb = zb[i];
    c.G = mpm_G<double>(c.h, c.hu, c.hv, m0[i] != 0, sd);
    return c;
}

template <typename T>
__device__ __forceinline__ ExnerCell<T> ec_lift(const ExnerCell<double> &c) {
    ExnerCell<T> r;
    r.h = T(c.h); r.hu = T(c.hu); r.hv = T(c.hv); r.zb = T(c.zb); r.G = T(c.G);
    return r;
}

__global__ __launch_bounds__(ADJ_BX *ADJ_BY) void k_adj_exner(F<double> hyd, const double *__restrict__ zb_in,
                                                              const uint8_t *__restrict__ m0,
                                                              const double *__restrict__ dtp, int ny, int nx,
                                                              double dx, pxh_sediments sd, F<double> lam_out,
                                                              F<double> lam_hyd, double *lam_zb_in,
                                                              double *__restrict__ part_dt,
                                                              double *__restrict__ part_par, int want_par) {
    const int c = blockIdx.x * ADJ_BX + threadIdx.x;
    const int r = blockIdx.y * ADJ_BY + threadIdx.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double ldt = 0.0, lp[4] = {0.0, 0.0, 0.0, 0.0};
    const bool par = want_par && sd.n_sed >= 1;
    if (r < ny && c < nx) {
        const int i = r * nx + c;
        const double dtdx = *dtp / dx;
        const bool mme = m0[i] != 0;
        const ExnerCell<double> me = ec_at(hyd, zb_in, sd, m0, i);
        ExnerCell<D4> md;  // own cell seeded: h, hu, hv, z_b
        md.h = D4::var(me.h, 0); md.hu = D4::var(me.hu, 1); md.hv = D4::var(me.hv, 2); md.zb = D4::var(me.zb, 3);
        md.G = mpm_G<D4>(md.h, md.hu, md.hv, mme, sd);
        D4 acc;
        const double lG = lam_out.G[i];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc.d[k] += lG * md.G.d[k];
        double div = 0.0;
        const D4 Ap = D4::var(sd.A[0], 0), Bp = D4::var(sd.B[0], 1), Cp = D4::var(sd.C[0], 2);
        const D4 Op = D4::var(sd.one_minus_por, 3);
        ExnerCell<D4> mp;  // own cell as a function of the parameters
        if (par) {
            mp = ec_lift<D4>(me);
            mp.G = mpm_G_par<D4>(mp.h, mp.hu, mp.hv, mme, sd, Ap, Bp, Cp, Op);
#pragma unroll
            for (int k = 0; k < 4; ++k) lp[k] += lG * mp.G.d[k];
        }
        auto edge = [&](auto ax, int j, bool own_left) {
            constexpr int AX = decltype(ax)::value;
            const ExnerCell<double> o = ec_at(hyd, zb_in, sd, m0, j);
            const bool mo = m0[j] != 0;
            const bool act = own_left ? exner_active(me, o, mme, mo) : exner_active(o, me, mo, mme);
            if (!act) return;
            const ExnerCell<D4> od = ec_lift<D4>(o);
            const D4 f = own_left ? exner_edge<AX, D4>(md, od) : exner_edge<AX, D4>(od, md);
            // zb'[L] -= dtdx f and zb'[R] += dtdx f
            const int L = own_left ? i : j, R = own_left ? j : i;
            const double coef = -dtdx * (lam_out.z_b[L] - lam_out.z_b[R]);
#pragma unroll
            for (int k = 0; k < 4; ++k) acc.d[k] += coef * f.d[k];
            div += own_left ? f.v : -f.v;
            if (par && own_left) {  // each edge's parameter derivative once, from its left cell
                ExnerCell<D4> op = od;
                op.G = mpm_G_par<D4>(op.h, op.hu, op.hv, mo, sd, Ap, Bp, Cp, Op);
                const D4 fp = exner_edge<AX, D4>(mp, op);
#pragma unroll
                for (int k = 0; k < 4; ++k) lp[k] += coef * fp.d[k];
            }
        };
        if (c > 0) edge(AxX{}, i - 1, false);
        if (c + 1 < nx) edge(AxX{}, i + 1, true);
        if (r > 0) edge(AxY{}, i - nx, false);
        if (r + 1 < ny) edge(AxY{}, i + nx, true);
        lam_hyd.h[i] += acc.d[0];
        lam_hyd.hu[i] += acc.d[1];
        lam_hyd.hv[i] += acc.d[2];
        lam_zb_in[i] += lam_out.z_b[i] + acc.d[3];
        ldt = lam_out.z_b[i] * (-div / dx);
    }
    block_sum_to(ldt, part_dt + blk);
    if (want_par) {
#pragma unroll
        for (int k = 0; k < 4; ++k) block_sum_to(lp[k], part_par + 4 * long(blk) + k);
    }
}

// Momentum corrections transposed, in place on lam hu, hv: zero where the
// forward zeroed the momentum.  h1: the hydro update before the fill; z, zb:
// the stage input's beds.
__global__ __launch_bounds__(ADJ_BX *ADJ_BY) void k_adj_momcorr(const double *__restrict__ h1,
                                                                const double *__restrict__ z,
                                                                const double *__restrict__ zb,
                                                                const uint8_t *__restrict__ m0, int ny, int nx,
                                                                F<double> lam) {
    const int c = blockIdx.x * ADJ_BX + threadIdx.x;
    const int r = blockIdx.y * ADJ_BY + threadIdx.y;
    if (r < 1 || c < 1 || r >= ny - 1 || c >= nx - 1) return;
    const int i = r * nx + c;
    bool keep_u, keep_v;
    momcorr_keep<double>(h1, z, zb, m0, nx, i, keep_u, keep_v);
    if (!keep_u) lam.hu[i] = 0.0;
    if (!keep_v) lam.hv[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Hydro update transposed.  in: the stage input; lam1: the adjoint of the
// hydro update (h1, hu1, hv1).  Adds into lam_in (h, hu, hv, z_b, and n_b: the
// Manning coefficient of the bed, when set) and the per-block d/d dt partials.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ HydroCell<T> hc_lift(const HydroCell<double> &c) {
    HydroCell<T> r;
    r.h = T(c.h); r.hu = T(c.hu); r.hv = T(c.hv); r.zb = T(c.zb); r.z = T(c.z); r.n = T(c.n);
    return r;
}

// Two wavefronts per SIMD (the dual-number edges want 270 registers, one
// wavefront's worth; held to 256 the kernel spills 20 and runs 5% faster).
__global__ __launch_bounds__(ADJ_BX *ADJ_BY) __attribute__((amdgpu_waves_per_eu(2))) void k_adj_hydro(
    F<double> in, const uint8_t *__restrict__ m0, const double *__restrict__ dtp, int ny, int nx, double dx,
    F<double> lam1, F<double> lam_in, double *__restrict__ part_dt) {
    const int c = blockIdx.x * ADJ_BX + threadIdx.x;
    const int r = blockIdx.y * ADJ_BY + threadIdx.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double ldt = 0.0;
    if (r < ny && c < nx) {
        const int i = r * nx + c;
        const double T_ = DRY_TOL;
        const double dtdx = *dtp / dx;
        const bool mme = m0[i] != 0;
        const HydroCell<double> me = load_hc(in, i);
        HydroCell<D5> md = hc_lift<D5>(me);
        md.h = D5::var(me.h, 0); md.hu = D5::var(me.hu, 1); md.hv = D5::var(me.hv, 2); md.zb = D5::var(me.zb, 3);
        md.n = D5::var(me.n, 4);
        D5 acc;
        double F3[3] = {0.0, 0.0, 0.0};  // this cell's flux sum, for d/d dt
        auto edge = [&](auto ax, int j, bool own_left) {
            constexpr int AX = decltype(ax)::value;
            const HydroCell<double> o = load_hc(in, j);
            if (!(((me.h > T_) || (o.h > T_)) && !(mme || m0[j] != 0))) return;
            const HydroCell<D5> od = hc_lift<D5>(o);
            D5 P[3], M[3];
            if (own_left) roe_exner_edge<AX, D5>(md, od, D5(dx), P, M);
            else roe_exner_edge<AX, D5>(od, md, D5(dx), P, M);
            // q1[L] -= dtdx M, q1[R] -= dtdx P
            const int L = own_left ? i : j, R = own_left ? j : i;
            const double lL[3] = {lam1.h[L], lam1.hu[L], lam1.hv[L]};
            const double lR[3] = {lam1.h[R], lam1.hu[R], lam1.hv[R]};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
#pragma unroll
                for (int k = 0; k < 5; ++k) acc.d[k] += -dtdx * (lL[q] * M[q].d[k] + lR[q] * P[q].d[k]);
                F3[q] += own_left ? M[q].v : P[q].v;
            }
        };
        if (c + 1 < nx) edge(AxX{}, i + 1, true);
        if (c > 0) edge(AxX{}, i - 1, false);
        if (r + 1 < ny) edge(AxY{}, i + nx, true);
        if (r > 0) edge(AxY{}, i - nx, false);
        const double l0 = lam1.h[i], l1 = lam1.hu[i], l2 = lam1.hv[i];
        lam_in.h[i] += l0 + acc.d[0];
        lam_in.hu[i] += l1 + acc.d[1];
        lam_in.hv[i] += l2 + acc.d[2];
        lam_in.z_b[i] += acc.d[3];
        if (lam_in.n_b) lam_in.n_b[i] += acc.d[4];
        ldt = -(l0 * F3[0] + l1 * F3[1] + l2 * F3[2]) / dx;
    }
    block_sum_to(ldt, part_dt + blk);
}

// compute_G of the initial state transposed: lam_s (h, hu, hv) += lamG dG,
// parameter partials (4 per block) with want_par; and the per-block sums of
// lam n_b over the cells compute_n gave the bed coefficient (z_b > SED_TOL).
__global__ __launch_bounds__(ADJ_BX *ADJ_BY) void k_adj_init(F<double> s, const uint8_t *__restrict__ m0, int ny,
                                                             int nx, pxh_sediments sd, F<double> lam,
                                                             double *__restrict__ part_par, int want_par) {
    const int c = blockIdx.x * ADJ_BX + threadIdx.x;
    const int r = blockIdx.y * ADJ_BY + threadIdx.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double lp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (r < ny && c < nx) {
        const int i = r * nx + c;
        const bool nod = m0[i] != 0;
        const double lG = lam.G[i];
        const D4 G = mpm_G<D4>(D4::var(s.h[i], 0), D4::var(s.hu[i], 1), D4::var(s.hv[i], 2), nod, sd);
        lam.h[i] += lG * G.d[0];
        lam.hu[i] += lG * G.d[1];
        lam.hv[i] += lG * G.d[2];
        if (want_par && sd.n_sed >= 1) {
            const D4 Gp = mpm_G_par<D4>(D4(s.h[i]), D4(s.hu[i]), D4(s.hv[i]), nod, sd, D4::var(sd.A[0], 0),
                                        D4::var(sd.B[0], 1), D4::var(sd.C[0], 2), D4::var(sd.one_minus_por, 3));
#pragma unroll
            for (int k = 0; k < 4; ++k) lp[k] = lG * Gp.d[k];
            if (s.z_b[i] > SED_TOL) lp[4] = lam.n_b[i];
        }
    }
    if (want_par) {
#pragma unroll
        for (int k = 0; k < 5; ++k) block_sum_to(lp[k], part_par + ADJ_NPAR * long(blk) + k);
    }
}

// ---------------------------------------------------------------------------
// The CFL dt transposed.  dt = max(cfl min_e dt_e, 1e-6): the derivative goes
// to the edge attaining the minimum, and is zero while the clamp holds.
// dt_e of the coupled cubic: the extreme roots t of the trisection equation
// 4 t^3 - 3 t = x (x the depressed cubic's scaled argument) move with x as
// dt/dx = 1 / (12 t^2 - 3), implicit differentiation of the equation; in the
// clip range |x| >= 1 the coded formula holds t fixed, so the derivative there
// is zero (roe_exner.py:185).
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void get_lambda_d(Dual<N> ut, Dual<N> at, Dual<N> ct, Dual<N> &lmin, Dual<N> &lmax) {
    using T = Dual<N>;
    if (at < 1e-5) {
        lmin = ut - ct;
        lmax = ut + ct;
        return;
    }
    const double a = -1.0;
    T b = -2.0 * ut;
    T c = (1.0 + at) * (ct * ct) - ut * ut;
    T d = ((-(ct * ct)) * at) * ut;
    T p = ((3.0 * a) * c - b * b) / 3.0;
    T q = ((2.0 * cube(b) - ((9.0 * a) * b) * c) + 27.0 * d) / -27.0;
    T x = (3.0 * q) / (2.0 * p) * d_sqrt(-3.0 / p);
    double tmin_v, tmax_v;
    trisect(x.v, tmin_v, tmax_v);
    T tmax(tmax_v), tmin(tmin_v);
    if (x.v < 1.0 && x.v > -1.0) {
        tmax = dual_unary(x, tmax_v, 1.0 / (12.0 * tmax_v * tmax_v - 3.0));
        tmin = dual_unary(x, tmin_v, 1.0 / (12.0 * tmin_v * tmin_v - 3.0));
    }
    T st = 2.0 * d_sqrt(-p / 3.0);
    T sh = -b / (3.0 * a);
    lmax = st * tmax + sh;
    lmin = st * tmin + sh;
}

// dt_edge (roe_exner.py:211-255) on Dual values.
template <int AX, int N>
__device__ __forceinline__ Dual<N> dt_edge_d(Dual<N> Ih, Dual<N> Ihu, Dual<N> Ihv, Dual<N> IG, Dual<N> Jh,
                                             Dual<N> Jhu, Dual<N> Jhv, Dual<N> JG, double dx) {
    using T = Dual<N>;
    const double T_ = DRY_TOL, g = GRAV;
    T hi = pmax(Ih, T(T_)), hj = pmax(Jh, T(T_));
    T si = d_sqrt(hi), sj = d_sqrt(hj);
    T den = si + sj;
    T ui = Ihu / hi, vi = Ihv / hi, uj = Jhu / hj, vj = Jhv / hj;
    T uhi, vhi, uhj, vhj;
    rot<AX>(ui, vi, uhi, vhi);
    rot<AX>(uj, vj, uhj, vhj);
    T ht = 0.5 * (hi + hj);
    T ut = (uhi * si + uhj * sj) / den;
    T ct = d_sqrt(g * ht);
    T gt = 0.5 * (IG + JG);
    T at = gt * (sq(uhi) + uhi * uhj + sq(uhj) + vhi * vhj) / (si * sj);
    T l1 = ut - ct, l3 = ut + ct;
    T lmin, lmax;
    get_lambda_d<N>(ut, at, ct, lmin, lmax);
    T rs = pmax(d_abs(lmin), d_abs(lmax));
    T ss = pmax(d_abs(l1), d_abs(l3));
    return (rs < VEL_TOL) ? T(dx) / ss : T(dx) / rs;
}

__device__ __forceinline__ bool dt_active(double hi, double hj, bool mi, bool mj) {
    return ((hi > DRY_TOL) || (hj > DRY_TOL)) && !(mi || mj);
}

// (value, edge) pairs: the smaller value wins, ties to the smaller edge id;
// id -1 marks no edge.
__device__ __forceinline__ void argmin_merge(double &v, long &id, double v2, long i2) {
    if (i2 < 0) return;
    if (id < 0 || v2 < v || (v2 == v && i2 < id)) {
        v = v2;
        id = i2;
    }
}

__device__ __forceinline__ void block_argmin(double &v, long &id) {
    __shared__ double rv[256];
    __shared__ long ri[256];
    const int t = threadIdx.x;
    rv[t] = v;
    ri[t] = id;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            double a = rv[t];
            long ia = ri[t];
            argmin_merge(a, ia, rv[t + w], ri[t + w]);
            rv[t] = a;
            ri[t] = ia;
        }
        __syncthreads();
    }
    v = rv[0];
    id = ri[0];
}

// Per-block minimum edge dt and its edge id (2 * cell + 0 for the east edge,
// + 1 for the south edge), evaluated exactly as the forward dt does.
__global__ __launch_bounds__(256) void k_adj_dt_argmin(F<double> s, const uint8_t *__restrict__ m0, int ny, int nx,
                                                      double dx, double *__restrict__ part_v,
                                                      long *__restrict__ part_i) {
    const long i = long(blockIdx.x) * 256 + threadIdx.x;
    double best = 0.0;
    long id = -1;
    if (i < long(ny) * nx) {
        const int r = int(i / nx), c = int(i - long(r) * nx);
        const bool mi = m0[i] != 0;
        if (c + 1 < nx && dt_active(s.h[i], s.h[i + 1], mi, m0[i + 1] != 0))
            argmin_merge(best, id,
                         dt_edge<0, double>(s.h[i], s.hu[i], s.hv[i], s.G[i], s.h[i + 1], s.hu[i + 1], s.hv[i + 1],
                                            s.G[i + 1], dx),
                         2 * i);
        if (r + 1 < ny && dt_active(s.h[i], s.h[i + nx], mi, m0[i + nx] != 0))
            argmin_merge(best, id,
                         dt_edge<1, double>(s.h[i], s.hu[i], s.hv[i], s.G[i], s.h[i + nx], s.hu[i + nx],
                                            s.hv[i + nx], s.G[i + nx], dx),
                         2 * i + 1);
    }
    block_argmin(best, id);
    if (threadIdx.x == 0) {
        part_v[blockIdx.x] = best;
        part_i[blockIdx.x] = id;
    }
}

// One block: the argmin over the partials, then lam_s (h, hu, hv, G of the
// edge's two cells) += lam_dt cfl d dt_e.  lam_dt: the summed partials.
__global__ __launch_bounds__(256) void k_adj_dt_apply(F<double> s, int nx, double dx, double cfl,
                                                     const double *__restrict__ part_v,
                                                     const long *__restrict__ part_i, int nblk,
                                                     const double *__restrict__ lam_dt, F<double> lam_s) {
    double best = 0.0;
    long id = -1;
    for (int k = threadIdx.x; k < nblk; k += 256) argmin_merge(best, id, part_v[k], part_i[k]);
    block_argmin(best, id);
    if (threadIdx.x != 0 || id < 0 || !(cfl * best > 1e-6)) return;  // all dry, or the clamp holds dt
    const long i = id / 2;
    const long j = (id & 1) ? i + nx : i + 1;
    const D8 Ih = D8::var(s.h[i], 0), Ihu = D8::var(s.hu[i], 1), Ihv = D8::var(s.hv[i], 2), IG = D8::var(s.G[i], 3);
    const D8 Jh = D8::var(s.h[j], 4), Jhu = D8::var(s.hu[j], 5), Jhv = D8::var(s.hv[j], 6), JG = D8::var(s.G[j], 7);
    const D8 v = (id & 1) ? dt_edge_d<1, 8>(Ih, Ihu, Ihv, IG, Jh, Jhu, Jhv, JG, dx)
                          : dt_edge_d<0, 8>(Ih, Ihu, Ihv, IG, Jh, Jhu, Jhv, JG, dx);
    const double w = (*lam_dt) * cfl;
    lam_s.h[i] += w * v.d[0]; lam_s.hu[i] += w * v.d[1]; lam_s.hv[i] += w * v.d[2]; lam_s.G[i] += w * v.d[3];
    lam_s.h[j] += w * v.d[4]; lam_s.hu[j] += w * v.d[5]; lam_s.hv[j] += w * v.d[6]; lam_s.G[j] += w * v.d[7];
}

// out[k] (+)= sum over nseg segments s and n partials of
// part[s * seg + j * stride + k0 + k]; one block per k.  accumulate = 0
// overwrites.
__global__ __launch_bounds__(256) void k_adj_sum(const double *__restrict__ part, long n, int stride, int k0,
                                                 double *__restrict__ out, int accumulate, int nseg = 1,
                                                 long seg = 0) {
    __shared__ double red[256];
    const int k = blockIdx.x;
    double v = 0.0;
    for (int q = 0; q < nseg; ++q)
        for (long j = threadIdx.x; j < n; j += 256) v += part[q * seg + j * stride + k0 + k];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = accumulate ? out[k] + red[0] : red[0];
}

// SSPRK2 combine a' = 0.5 a + 0.5 c transposed: lam_c = 0.5 lam_a',
// lam_a = 0.5 lam_a' (h, hu, hv, z_b, G); also lam_b = 0 (h, hu, hv, z_b, G,
// n_b unused) for the stage-2 adjoint to accumulate into.
__global__ __launch_bounds__(256) void k_adj_split(F<double> la, F<double> lc, F<double> lb, long ncell) {
    const long i = long(blockIdx.x) * 256 + threadIdx.x;
    if (i >= ncell) return;
    double *fa[5] = {la.h, la.hu, la.hv, la.z_b, la.G};
    double *fc[5] = {lc.h, lc.hu, lc.hv, lc.z_b, lc.G};
    double *fb[5] = {lb.h, lb.hu, lb.hv, lb.z_b, lb.G};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const double v = 0.5 * fa[q][i];
        fa[q][i] = v;
        fc[q][i] = v;
        fb[q][i] = 0.0;
    }
}

}  // namespace pxh
