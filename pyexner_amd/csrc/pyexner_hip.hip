// PyExner per-step hot path for gfx950: kernels and the extern "C" entry points
// declared in include/pyexner_hip.h.
//
// Layout: structure-of-arrays (ny, nx) fields in HBM, one uint8 nodata mask.
// A Forward-Euler step (solvers/roe_exner_solver.py:80-116) runs as
//   hydro flux + update -> momentum corrections -> ring fill (h, hu, hv)
//   -> G -> Exner flux + bed update -> ring fill (z_b, G)
// and the CFL dt is a two-level min reduction into a device scalar.
#include <initializer_list>
#include <mutex>

#include "common.hpp"
#include "flux_roe.hpp"
#include "flux_roe_exner.hpp"
#include "cell.hpp"
#include "strip.hpp"
#include "stage.hpp"
#include "halo.hpp"
#include "ensemble.hpp"

namespace pxh {

constexpr int BX = 64, BY = 4;

static inline dim3 grid2d(int ny, int nx) {
    return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY);
}

// ---------------------------------------------------------------------------
// Roe-Exner hydro: roe_solve_2D (roe_exner.py:535-612), one thread per cell.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(BX *BY) void k_rx_hydro(F<T> in, F<T> out, const uint8_t *__restrict__ m0,
                                                      const T *__restrict__ dtp, int ny, int nx, T dx) {
    const int c = blockIdx.x * BX + threadIdx.x;
    const int r = blockIdx.y * BY + threadIdx.y;
    if (r >= ny || c >= nx) return;
    const int i = r * nx + c;
    T h1, hu1, hv1;
    rx_hydro_cell<T>(in, m0, *dtp, ny, nx, dx, r, c, h1, hu1, hv1);
    out.h[i] = h1;
    out.hu[i] = hu1;
    out.hv[i] = hv1;
}

// momentum_corrections (roe_exner.py:14-62) on interior cells, in place on the
// updated hu/hv; z + z_b uses the bed before this step's Exner update.
template <typename T>
__global__ __launch_bounds__(BX *BY) void k_rx_momcorr(F<T> in, F<T> out, const uint8_t *__restrict__ m0,
                                                        int ny, int nx) {
    const int c = blockIdx.x * BX + threadIdx.x;
    const int r = blockIdx.y * BY + threadIdx.y;
    if (r < 1 || c < 1 || r >= ny - 1 || c >= nx - 1) return;
    const int i = r * nx + c;
    bool keep_u, keep_v;
    momcorr_keep<T>(out.h, in.z, in.z_b, m0, nx, i, keep_u, keep_v);
    if (!keep_u) out.hu[i] = T(0);
    if (!keep_v) out.hv[i] = T(0);
}

// G for every cell from the post-fill hydro state (roe_exner_solver.py:99).
template <typename T>
__global__ __launch_bounds__(256) void k_rx_G(F<T> in, F<T> out, const uint8_t *__restrict__ m0, int ncell,
                                               pxh_sediments sd) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ncell) return;
    const bool nod = m0[i] != 0;
    out.G[i] = sd.mpm ? mpm_G<T>(out.h[i], out.hu[i], out.hv[i], nod, sd) : const_G<T>(out.h[i], in.G[i], nod);
}

// exner_solve_2D (roe_exner.py:665-717): reads the updated h, hu, hv, G and
// the bed of the step start; div = ((0 - Fx[c-1]) + Fx[c]) - Fy[r-1] + Fy[r].
template <typename T>
__global__ __launch_bounds__(BX *BY) void k_rx_exner(F<T> in, F<T> out, const uint8_t *__restrict__ m0,
                                                      const T *__restrict__ dtp, int ny, int nx, T dx) {
    const int c = blockIdx.x * BX + threadIdx.x;
    const int r = blockIdx.y * BY + threadIdx.y;
    if (r >= ny || c >= nx) return;
    const int i = r * nx + c;
    const ExnerCell<T> me = load_ec(out, in.z_b, i);
    const bool mme = m0[i] != 0;
    T div = T(0);
    if (c > 0) {
        const int j = i - 1;
        const ExnerCell<T> o = load_ec(out, in.z_b, j);
        const T f = exner_active(o, me, m0[j] != 0, mme) ? exner_edge<0>(o, me) : T(0);
        div = div + (-f);
    }
    if (c + 1 < nx) {
        const int j = i + 1;
        const ExnerCell<T> o = load_ec(out, in.z_b, j);
        const T f = exner_active(me, o, mme, m0[j] != 0) ? exner_edge<0>(me, o) : T(0);
        div = div + f;
    }
    if (r > 0) {
        const int j = i - nx;
        const ExnerCell<T> o = load_ec(out, in.z_b, j);
        const T f = exner_active(o, me, m0[j] != 0, mme) ? exner_edge<1>(o, me) : T(0);
        div = div + (-f);
    }
    if (r + 1 < ny) {
        const int j = i + nx;
        const ExnerCell<T> o = load_ec(out, in.z_b, j);
        const T f = exner_active(me, o, mme, m0[j] != 0) ? exner_edge<1>(me, o) : T(0);
        div = div + f;
    }
    out.z_b[i] = me.zb - *dtp * div / dx;
}

// ---------------------------------------------------------------------------
// Plain Roe hydro (roe.py:251-315): no active-edge masking.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ RoeCell<T> load_rc(const F<T> &s, int i) {
    RoeCell<T> c;
    c.h = s.h[i]; c.hu = s.hu[i]; c.hv = s.hv[i]; c.z = s.z[i]; c.n = s.n[i];
    return c;
}

template <typename T>
__global__ __launch_bounds__(BX *BY) void k_roe_hydro(F<T> in, F<T> out, const T *__restrict__ dtp, int ny,
                                                       int nx, T dx) {
    const int c = blockIdx.x * BX + threadIdx.x;
    const int r = blockIdx.y * BY + threadIdx.y;
    if (r >= ny || c >= nx) return;
    const int i = r * nx + c;
    const RoeCell<T> me = load_rc(in, i);
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    T P[3], M[3];
    if (c + 1 < nx) {
        roe_edge<0>(me, load_rc(in, i + 1), dx, P, M);
        acc0 = acc0 + M[0]; acc1 = acc1 + M[1]; acc2 = acc2 + M[2];
    }
    if (c > 0) {
        roe_edge<0>(load_rc(in, i - 1), me, dx, P, M);
        acc0 = acc0 + P[0]; acc1 = acc1 + P[1]; acc2 = acc2 + P[2];
    }
    if (r + 1 < ny) {
        roe_edge<1>(me, load_rc(in, i + nx), dx, P, M);
        acc0 = acc0 + M[0]; acc1 = acc1 + M[1]; acc2 = acc2 + M[2];
    }
    if (r > 0) {
        roe_edge<1>(load_rc(in, i - nx), me, dx, P, M);
        acc0 = acc0 + P[0]; acc1 = acc1 + P[1]; acc2 = acc2 + P[2];
    }
    const T dt = *dtp;
    out.h[i] = me.h - dt * acc0 / dx;
    out.hu[i] = me.hu - dt * acc1 / dx;
    out.hv[i] = me.hv - dt * acc2 / dx;
}

// ---------------------------------------------------------------------------
// CFL dt: per-block NaN-propagating min into partials, then a one-block final
// reduction that applies cfl and the 1e-6 clamp (roe_exner_solver.py:118-122).
// ---------------------------------------------------------------------------
constexpr int DT_THREADS = 256;
constexpr int DT_MAX_BLOCKS = 4096;

template <typename T>
__device__ __forceinline__ T block_min(T v) {
    __shared__ T red[DT_THREADS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < DT_THREADS / 64; ++k) v = pmin(v, red[k]);
    }
    return v;
}

template <typename T>
__global__ __launch_bounds__(DT_THREADS) void k_rx_dt_partial(F<T> s, const uint8_t *__restrict__ m0, int ny,
                                                              int nx, T dx, T *__restrict__ partial) {
    const T T_ = C<T>(DRY_TOL);
    T v = inf_v<T>();
    const long ncell = long(ny) * nx;
    for (long i = long(blockIdx.x) * DT_THREADS + threadIdx.x; i < ncell; i += long(gridDim.x) * DT_THREADS) {
        const int r = int(i / nx), c = int(i - long(r) * nx);
        const T hi = s.h[i];
        const bool mi = m0[i] != 0;
        if (c + 1 < nx) {
            const long j = i + 1;
            const T hj = s.h[j];
            if (((hi > T_) || (hj > T_)) && !(mi || m0[j]))
                v = pmin(v, dt_edge<0>(hi, s.hu[i], s.hv[i], s.G[i], hj, s.hu[j], s.hv[j], s.G[j], dx));
        }
        if (r + 1 < ny) {
            const long j = i + nx;
            const T hj = s.h[j];
            if (((hi > T_) || (hj > T_)) && !(mi || m0[j]))
                v = pmin(v, dt_edge<1>(hi, s.hu[i], s.hv[i], s.G[i], hj, s.hu[j], s.hv[j], s.G[j], dx));
        }
    }
    v = block_min(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// roe.py:13-30: cell-based speeds over every cell (the mask has no effect).
template <typename T>
__global__ __launch_bounds__(DT_THREADS) void k_roe_dt_partial(F<T> s, long ncell, T dx, T *__restrict__ partial) {
    const T T_ = C<T>(DRY_TOL), g = C<T>(GRAV), V = C<T>(VEL_TOL), zero = T(0);
    T v = inf_v<T>();
    for (long i = long(blockIdx.x) * DT_THREADS + threadIdx.x; i < ncell; i += long(gridDim.x) * DT_THREADS) {
        const T hr = s.h[i];
        const bool wet = hr > T_;
        const T h = wet ? hr : zero;
        const T hv = wet ? s.hv[i] : zero;
        const T hu = wet ? s.hu[i] : zero;
        const T u = hu / (h + T_), vv = hv / (h + T_);
        const T cc = d_sqrt(g * h);
        const T dtx = dx / (d_abs(u) + cc + V);
        const T dty = dx / (d_abs(vv) + cc + V);
        v = pmin(v, pmin(dtx, dty));
    }
    v = block_min(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// The final min over the partials (tens of thousands for the strip sweeps):
// one block of 1024 threads with eight loads in flight per thread.  min is
// exact, so the order of the reduction does not change the result (a NaN
// anywhere still wins through pmin).
constexpr int FIN_THREADS = 1024;
constexpr int FIN_UNROLL = 8;

template <typename T>
__global__ __launch_bounds__(FIN_THREADS) void k_dt_final(const T *__restrict__ partial, int nblk, T cfl, int clamp,
                                                          T *__restrict__ out) {
    __shared__ T red[FIN_THREADS / 64];
    T acc[FIN_UNROLL];
#pragma unroll
    for (int u = 0; u < FIN_UNROLL; ++u) acc[u] = inf_v<T>();
    int k = threadIdx.x;
    for (; k + (FIN_UNROLL - 1) * FIN_THREADS < nblk; k += FIN_UNROLL * FIN_THREADS) {
#pragma unroll
        for (int u = 0; u < FIN_UNROLL; ++u) acc[u] = pmin(acc[u], partial[k + u * FIN_THREADS]);
    }
    for (; k < nblk; k += FIN_THREADS) acc[0] = pmin(acc[0], partial[k]);
    T v = acc[0];
#pragma unroll
    for (int u = 1; u < FIN_UNROLL; ++u) v = pmin(v, acc[u]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < FIN_THREADS / 64; ++w) v = pmin(v, red[w]);
        T d = cfl * v;
        if (clamp) d = pmax(d, C<T>(1e-6));
        *out = d;
    }
}

// ---------------------------------------------------------------------------
// Non-finite scan: the failure flag an output window reads back
// ---------------------------------------------------------------------------
constexpr int NF_THREADS = 256;
constexpr int NF_MAX_BLOCKS = 4096;

template <typename T>
struct NfSet {
    const T *p[6];
    int nf;
};

__global__ void k_zero_u64(unsigned long long *x) { *x = 0ull; }

// Cells with a NaN or infinite value in any of the set's fields; one atomic
// per block that found any.
template <typename T>
__global__ __launch_bounds__(NF_THREADS) void k_count_nonfinite(NfSet<T> s, long ncell,
                                                                unsigned long long *count) {
    __shared__ unsigned red[NF_THREADS / 64];
    unsigned cnt = 0;
    for (long i = long(blockIdx.x) * NF_THREADS + threadIdx.x; i < ncell; i += long(gridDim.x) * NF_THREADS) {
        bool bad = false;
        for (int f = 0; f < s.nf; ++f) bad |= !__builtin_isfinite(s.p[f][i]);
        cnt += bad ? 1u : 0u;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < NF_THREADS / 64; ++w) tot += red[w];
        if (tot) atomicAdd(count, tot);
    }
}

template <typename T>
static int count_nonfinite_t(const pxh_grid *g, const pxh_fields *s, unsigned long long *count, hipStream_t st) {
    NfSet<T> set{};
    const void *cand[6] = {s->h, s->hu, s->hv, s->z_b, sizeof(T) == 4 ? s->z_b_lo : nullptr, s->G};
    for (const void *q : cand)
        if (q) set.p[set.nf++] = static_cast<const T *>(q);
    const long n = long(g->ny) * g->nx;
    const long want = (n + NF_THREADS - 1) / NF_THREADS;
    const int nb = int(want < NF_MAX_BLOCKS ? want : NF_MAX_BLOCKS);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, count);
    hipLaunchKernelGGL(k_count_nonfinite<T>, dim3(nb), dim3(NF_THREADS), 0, st, set, n, count);
    return PXH_OK;
}

// ---------------------------------------------------------------------------
// SSPRK2 combine and init
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_combine(F<T> a, F<T> b, F<T> o, long ncell) {
    const long i = long(blockIdx.x) * 256 + threadIdx.x;
    if (i >= ncell) return;
    const T hf = T(0.5);
    o.h[i] = hf * a.h[i] + hf * b.h[i];
    o.hu[i] = hf * a.hu[i] + hf * b.hu[i];
    o.hv[i] = hf * a.hv[i] + hf * b.hv[i];
    T zc = hf * a.z_b[i] + hf * b.z_b[i];
    if constexpr (sizeof(T) == 4) {
        if (o.z_b_lo) {
            T zl;
            cb_half_sum(a.z_b[i], a.z_b_lo[i], b.z_b[i], b.z_b_lo[i], zc, zl);
            o.z_b_lo[i] = zl;
        }
    }
    o.z_b[i] = zc;
    o.G[i] = hf * a.G[i] + hf * b.G[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_rx_init(F<T> s, const uint8_t *__restrict__ m0, long ncell,
                                                  pxh_sediments sd) {
    const long i = long(blockIdx.x) * 256 + threadIdx.x;
    if (i >= ncell) return;
    const bool nod = m0[i] != 0;
    if (sd.mpm) {
        // compute_n (roe_exner.py:64-73)
        s.n_b[i] = (s.z_b[i] > C<T>(SED_TOL)) ? C<T>(sd.nb_sed) : s.n[i];
        s.G[i] = mpm_G<T>(s.h[i], s.hu[i], s.hv[i], nod, sd);
    } else {
        s.n_b[i] = s.n[i];
        s.G[i] = const_G<T>(s.h[i], s.G[i], nod);
    }
}

// ---------------------------------------------------------------------------
// Composed ring fill.  Each entry copies one field group from src to target;
// momentum entries apply the chain of handler transforms in order
// (roeexner_reflective.py:31-59, roe_reflective.py:34-39).
// ---------------------------------------------------------------------------
// The entries one ring fill visits: range r is entries [start[r], start[r] +
// off[r+1] - off[r]), taken by threads [off[r], off[r+1]).
struct BcSel {
    int n;
    int start[PXH_N_GROUPS];
    int off[PXH_N_GROUPS + 1];
};

template <typename T>
// sq: the element stride of s.h, hu, hv, z_b (4 where they are the packed
// records of PXH_PACKED_STAGE, else 1).
__global__ __launch_bounds__(256) void k_bc(F<T> s, const pxh_bc_entry *__restrict__ ent, BcSel sel, pxh_bc bc,
                                            int sq) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= sel.off[sel.n]) return;
    // the range holding k, with constant indices into sel (scalar registers,
    // no per-lane load of the table)
    int first = sel.start[0];
#pragma unroll
    for (int q = 1; q < PXH_N_GROUPS; ++q)
        if (q < sel.n && k >= sel.off[q]) first = sel.start[q] - sel.off[q];
    const BcEnt be = bc_ent(ent, first + k);
    const pxh_bc_entry e = be.first();
    if (!((bc.group_mask >> e.group) & 1u)) return;
    if (e.nchain & PXH_BC_VALUE) {
        bc_value_apply<T>(s, e, bc, sq);
        return;
    }
    const int t = e.target, src = e.src;
    const long tq = long(t) * sq, sqi = long(src) * sq;
    switch (e.group) {
        case PXH_G_H: s.h[tq] = s.h[sqi]; break;
        case PXH_G_Z: s.z[t] = s.z[src]; break;
        case PXH_G_ZB:
            s.z_b[tq] = s.z_b[sqi];
            if (s.z_b_lo) s.z_b_lo[t] = s.z_b_lo[src];
            break;
        case PXH_G_N: s.n[t] = s.n[src]; break;
        case PXH_G_NB: s.n_b[t] = s.n_b[src]; break;
        case PXH_G_G: s.G[t] = s.G[src]; break;
        case PXH_G_HUV: {
            T hu = s.hu[sqi], hv = s.hv[sqi];
            for (int q = 0; q < e.nchain; ++q) bc_transform<T>(bc.handlers[be.chain(q)], hu, hv);
            s.hu[tq] = hu;
            s.hv[tq] = hv;
        } break;
        default: break;
    }
}

// ---------------------------------------------------------------------------
// host-side launch helpers
// ---------------------------------------------------------------------------
static inline int dt_blocks(const pxh_grid *g) {
    long n = (long(g->ny) * g->nx + DT_THREADS - 1) / DT_THREADS;
    return int(n < DT_MAX_BLOCKS ? n : DT_MAX_BLOCKS);
}

// The dt workspace (pxh_dt_workspace_size): partial minima, 8 bytes each, for
// the largest launch sequence that reduces a dt, then the seam-column buffer
// of the combine sweeps (seam_rec).
static int64_t dt_part_slots(const pxh_grid *g) {
    const dim3 sa = strip_grid<SO_WIDE, SR_BED>(g->ny, g->nx), fg = stage_grid(g->ny, g->nx, strip_rows(g->flags));
    int64_t nb = int64_t(sa.x) * sa.y + seam_blocks(g->ny, g->nx);
    const int64_t nfused = int64_t(fg.x) * fg.y + (band_dt_work(g->ny, g->nx, strip_rows(g->flags)) + 255) / 256;
    if (nfused > nb) nb = nfused;
    nb = nb > DT_MAX_BLOCKS ? nb : DT_MAX_BLOCKS;
    return (nb + 31) & ~int64_t(31);  // the buffer after it starts 256-byte aligned
}

template <typename T>
static T *seam_cols(const pxh_grid *g, void *ws) {
    return reinterpret_cast<T *>(static_cast<char *>(ws) + dt_part_slots(g) * 8);
}

template <typename T>
// sq: the element stride of s.h, hu, hv, z_b (4 for the packed records of
// PXH_PACKED_STAGE).
static int bc_launch(const pxh_grid *g, const F<T> &s, const pxh_bc *bc, uint32_t groups, hipStream_t st,
                     int sq = 1) {
    if (!bc || bc->n_entries <= 0) return PXH_OK;
    pxh_bc b = *bc;
    b.group_mask &= groups;
    if (!b.group_mask) return PXH_OK;
    if (bc->n_levels > 1) {
        // one launch per dependency level, in order: each level reads only
        // values no earlier level wrote (k_bc skips the unselected groups)
        for (int L = 0; L < bc->n_levels; ++L) {
            BcSel sel{};
            sel.n = 1;
            sel.start[0] = bc->level_start[L];
            sel.off[1] = bc->level_start[L + 1] - bc->level_start[L];
            if (sel.off[1] <= 0) continue;
            hipLaunchKernelGGL(k_bc<T>, dim3((sel.off[1] + 255) / 256), dim3(256), 0, st, s, bc->entries, sel, b,
                               sq);
            if (int e = launch_status()) return e;
        }
        return PXH_OK;
    }
    BcSel sel{};
    if (bc->flags & PXH_BC_GROUPED) {
        for (int q = 0; q < PXH_N_GROUPS; ++q) {
            const int cnt = bc->group_start[q + 1] - bc->group_start[q];
            if (!((b.group_mask >> q) & 1u) || cnt <= 0) continue;
            sel.start[sel.n] = bc->group_start[q];
            sel.off[sel.n + 1] = sel.off[sel.n] + cnt;
            ++sel.n;
        }
        if (sel.n == 0) return PXH_OK;
    } else {
        sel.n = 1;
        sel.off[1] = bc->n_entries;
    }
    const int n = sel.off[sel.n];
    hipLaunchKernelGGL(k_bc<T>, dim3((n + 255) / 256), dim3(256), 0, st, s, bc->entries, sel, b, sq);
    return launch_status();
}

// The fused stage (stage.hpp) for frame-local boundary programs, as two
// launch groups the caller separates with the (h, hu, hv) ring fill of the
// stage's hydro result:
//   stage sweep, old == none: out.{h,hu,hv} of every cell, out.{z_b, z_b_lo
//     (, G when write_G)} of the cells off the two-cell frame;
//   stage sweep, old set: out.{h,hu,hv} of the cells within three cells of the
//     edge, old = 0.5 old + 0.5 S(in) off the frame, and with partial the
//     per-block dt minima of that combined state within each strip in
//     partial[0, nf);
//   frame sweep: the frame's bed update from the filled hyd and in's bed
//     (MODE 0 into out, MODE 1 combined into old), and with partial the dt
//     minima of the frame's edges and of the edges across strip seams, each
//     folded with a share of the stage sweep's minima, in partial[nf, nf +
//     nbd), followed by the final reduction of those nbd minima into *dtp.
template <typename T, bool FAST>
// packed: the first stage's output (out without old, in with it) keeps its
// h, hu, hv, z_b as the four-float records of PXH_PACKED_STAGE at its h.
static int stage_sweep_f(const pxh_grid *g, const F<T> &in, const F<T> &out, const F<T> *old, const uint8_t *m0,
                         const pxh_sediments *sd, const T *dtp, int write_G, T *part, hipStream_t st,
                         bool packed = false) {
    const int ny = g->ny, nx = g->nx;
    const T dx = static_cast<T>(g->dx);
    const F<T> none{};
    const int sr = strip_rows(g->flags);
    const dim3 fg = stage_grid(ny, nx, sr);
    const bool nn = (g->flags & PXH_NO_NODATA) != 0, zu = (g->flags & PXH_Z_UNIFORM) != 0;
    const int w1 = PXH_STAGE1_WAVES(g->flags), w2 = PXH_STAGE2_WAVES(g->flags);
    if (!old)
        launch_stage<T, FAST, 0, false>(fg, st, in, out, none, none, m0, dtp, ny, nx, sr, dx, sed_k(*sd), write_G,
                                        static_cast<T *>(nullptr), static_cast<T *>(nullptr), nn, w1, zu, packed);
    else if (part)
        launch_stage<T, FAST, 1, true>(fg, st, in, none, out, *old, m0, dtp, ny, nx, sr, dx, sed_k(*sd), 1, part,
                                       seam_cols<T>(g, part), nn, w2, zu, packed);
    else
        launch_stage<T, FAST, 1, false>(fg, st, in, none, out, *old, m0, dtp, ny, nx, sr, dx, sed_k(*sd), 1,
                                        static_cast<T *>(nullptr), static_cast<T *>(nullptr), nn, w2, zu, packed);
    return launch_status();
}

// hs / zs / os: the element strides of hyd's h, hu, hv / of bed.z_b / of
// out.z_b (4 where they are the packed records of PXH_PACKED_STAGE).
template <typename T, bool FAST>
static int frame_sweep_f(const pxh_grid *g, const F<T> &hyd, const F<T> &bed, const F<T> &out, const F<T> *old,
                         const uint8_t *m0, const pxh_sediments *sd, T *dtp, int write_G, double cfl, T *part,
                         hipStream_t st, int hs = 1, int zs = 1, int os = 1) {
    const int ny = g->ny, nx = g->nx;
    const T dx = static_cast<T>(g->dx);
    const F<T> none{};
    const long nband = frame_cells(ny, nx);
    const int nbb = int((nband + 255) / 256);  // >= 1: a non-empty grid has a frame
    if (!old) {
        hipLaunchKernelGGL((k_rx_band_bed<T, FAST, 0>), dim3(nbb), dim3(256), 0, st, hyd, bed, out, none, m0, dtp, ny,
                           nx, dx, sed_k(*sd), write_G, nband, hs, zs, os);
        return launch_status();
    }
    hipLaunchKernelGGL((k_rx_band_bed<T, FAST, 1>), dim3(nbb), dim3(256), 0, st, hyd, bed, none, *old, m0, dtp, ny, nx,
                       dx, sed_k(*sd), 1, nband, hs, zs, os);
    if (int e = launch_status()) return e;
    if (!part) return PXH_OK;
    const int sr = strip_rows(g->flags);
    const dim3 fg = stage_grid(ny, nx, sr);
    const int nf = int(fg.x * fg.y);
    const long nwork = band_dt_work(ny, nx, sr);
    const int nbd = int((nwork + 255) / 256);
    hipLaunchKernelGGL((k_rx_band_dt<T, FAST>), dim3(nbd), dim3(256), 0, st, *old, m0, seam_cols<T>(g, part), ny, nx,
                       dx, sr, nband, nwork, part, nf, part + nf);
    if (int e = launch_status()) return e;
    hipLaunchKernelGGL(k_dt_final<T>, dim3(1), dim3(FIN_THREADS), 0, st, part + nf, nbd, static_cast<T>(cfl), 1, dtp);
    return launch_status();
}

// One Forward-Euler step (roe_exner_solver.py:80-116) as two strip sweeps and
// two ring fills: hydro + momentum corrections -> fill (h, hu, hv) ->
// G + Exner -> fill (z_b, G).  With a frame-local boundary program: stage
// sweep -> fill (h, hu, hv) -> frame sweep; the caller fills (z_b, G) after.
// packed: out keeps h, hu, hv, z_b as the four-float records of
// PXH_PACKED_STAGE, out.hu / hv / z_b pointing at out.h + 1 / 2 / 3.
template <typename T, bool FAST>
static int rx_stage0_fused(const pxh_grid *g, const F<T> &in, const F<T> &out, const uint8_t *m0,
                           const pxh_sediments *sd, const pxh_bc *bc, const T *dtp, int write_G, hipStream_t st,
                           bool packed = false) {
    const int sq = packed ? 4 : 1;
    if (int e = stage_sweep_f<T, FAST>(g, in, out, nullptr, m0, sd, dtp, write_G, nullptr, st, packed)) return e;
    if (int e = bc_launch<T>(g, out, bc, (1u << PXH_G_H) | (1u << PXH_G_HUV), st, sq)) return e;
    return frame_sweep_f<T, FAST>(g, out, in, out, nullptr, m0, sd, const_cast<T *>(dtp), write_G, 0.0, nullptr, st,
                                  sq, 1, sq);
}

// The fused stage addresses a strip's rows with 32-bit byte offsets from the
// strip's first row: sr + 3 rows of nx elements must fit.
static bool stage_fits(const pxh_grid *g) {
    const long esz = g->dtype == PXH_F64 ? 8 : 4;
    return long(strip_rows(g->flags) + 4) * g->nx * esz < (1L << 32);
}

static bool frame_fused(const pxh_grid *g, const pxh_bc *bc) {
    return bc && (bc->flags & PXH_BC_FRAME) && stage_fits(g);
}

template <typename T, bool FAST>
static int rx_step_t(const pxh_grid *g, const pxh_fields *in_, const pxh_fields *out_, const uint8_t *m0,
                     const pxh_sediments *sd, const pxh_bc *bc, const void *dt, hipStream_t st) {
    F<T> in = cast_fields<T>(in_), out = cast_fields<T>(out_);
    const int ny = g->ny, nx = g->nx;
    const T dx = static_cast<T>(g->dx);
    const T *dtp = static_cast<const T *>(dt);
    if (frame_fused(g, bc)) {
        if (int e = rx_stage0_fused<T, FAST>(g, in, out, m0, sd, bc, dtp, 1, st)) return e;
        return bc_launch<T>(g, out, bc, (1u << PXH_G_ZB) | (1u << PXH_G_G), st);
    }
    const dim3 hg = strip_grid<SO_WIDE, SR_HYDRO>(ny, nx), bg = strip_grid<SO_WIDE, SR_BED>(ny, nx), sb(SW * SWPB);
    hipLaunchKernelGGL((k_rx_hydro_strip<T, FAST>), hg, sb, 0, st, in, out, m0, dtp, ny, nx, dx, 1);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(g, out, bc, (1u << PXH_G_H) | (1u << PXH_G_HUV), st)) return e;
    hipLaunchKernelGGL((k_rx_bed_strip<T, FAST, 0, false>), bg, sb, 0, st, out, in.z_b, in.z_b_lo, in.G, out, out,
                       m0, dtp, ny, nx, dx, sed_k(*sd), 1, nullptr, nullptr);
    if (int e = launch_status()) return e;
    return bc_launch<T>(g, out, bc, (1u << PXH_G_ZB) | (1u << PXH_G_G), st);
}

// The seam edges of a two-sweep combine (k_rx_seam_dt) into part[0, seam_blocks).
template <typename T, bool FAST>
static int seam_dt(const pxh_grid *g, const F<T> &s, const uint8_t *m0, const T *xs, T *part, hipStream_t st) {
    const int nsb = seam_blocks(g->ny, g->nx);
    if (nsb == 0) return PXH_OK;
    hipLaunchKernelGGL((k_rx_seam_dt<T, FAST>), dim3(nsb), dim3(256), 0, st, s, m0, xs, g->ny, g->nx,
                       static_cast<T>(g->dx), part);
    return launch_status();
}

// True when the program holds ring values that are not fixed points of the
// SSPRK2 average: a bed that moves with time (PXH_BV_AFFINE_T: Berthon,
// NormalFlowDepth) or an offset of a cell the step changes (PXH_BV_OFFSET:
// SteepFall).  The reference averages the ring like every other cell
// (integrators/SSPRK2.py:40), so such a ring after a step is
// 0.5 a_ring + 0.5 fill(c)_ring, not the fill of the averaged state.
static bool ring_values_averaged(const pxh_bc *bc) {
    if (!bc || bc->n_entries <= 0) return false;
    const int n = bc->n_values < PXH_MAX_VALUES ? bc->n_values : PXH_MAX_VALUES;
    for (int k = 0; k < n; ++k)
        if (bc->values[k].kind == PXH_BV_AFFINE_T || bc->values[k].kind == PXH_BV_OFFSET) return true;
    return false;
}

// One SSPRK2 step (integrators/SSPRK2.py:32-40) on three buffer sets:
// dt from a (unless dt_ready); b = S(a); c = S(b) with the combine
// a = 0.5 a + 0.5 c fused into the second bed sweep, which also reduces the
// CFL dt of the new a, so dt holds compute_dt_fn(a) on return.  In MPM mode
// the stage-1 G is never read, so it is not stored.  The fused dt reads G of
// ring cells before the final (z_b, G) fill; every handler kind here either
// copies G together with h, hu, hv from one source (Reflective, Periodic) or
// leaves it to compute_G of the filled h, hu, hv (Transmissive), and compute_G
// depends on the velocity only through |u|^2, so the fill cannot change it.
// The final (z_b, G) fill of the averaged state equals the average of the
// filled states only for copies and constants; a program with bed values
// that move (ring_values_averaged) takes two whole steps, each with its full
// fill, then the combine and the dt of the combined state.
template <typename T, bool FAST>
static int rx_ssprk2_t(const pxh_grid *g, const pxh_fields *a_, const pxh_fields *b_, const pxh_fields *c_,
                       const uint8_t *m0, const pxh_sediments *sd, const pxh_bc *bc, double cfl, void *dt, void *ws,
                       int dt_ready, hipStream_t st) {
    F<T> a = cast_fields<T>(a_), b = cast_fields<T>(b_), c = cast_fields<T>(c_);
    const int ny = g->ny, nx = g->nx;
    const T dx = static_cast<T>(g->dx);
    T *dtp = static_cast<T *>(dt);
    const dim3 hg = strip_grid<SO_WIDE, SR_HYDRO>(ny, nx), bg = strip_grid<SO_WIDE, SR_BED>(ny, nx), sb(SW * SWPB);
    const dim3 dg = strip_grid<SO_WIDE, SR_BED>(ny, nx);
    const int nb = int(bg.x * bg.y);  // partial minima of the combine sweep
    T *part = static_cast<T *>(ws);
    if (!dt_ready) {
        hipLaunchKernelGGL((k_rx_dt_strip<T, FAST>), dg, sb, 0, st, a, m0, ny, nx, dx, part);
        if (int e = launch_status()) return e;
        hipLaunchKernelGGL(k_dt_final<T>, dim3(1), dim3(FIN_THREADS), 0, st, part, int(dg.x * dg.y),
                           static_cast<T>(cfl), 1, dtp);
        if (int e = launch_status()) return e;
    }
    if (ring_values_averaged(bc)) {
        if (int e = rx_step_t<T, FAST>(g, a_, b_, m0, sd, bc, dtp, st)) return e;
        if (int e = rx_step_t<T, FAST>(g, b_, c_, m0, sd, bc, dtp, st)) return e;
        const long n = long(ny) * nx;
        hipLaunchKernelGGL(k_combine<T>, dim3((n + 255) / 256), dim3(256), 0, st, a, c, a, n);
        if (int e = launch_status()) return e;
        hipLaunchKernelGGL((k_rx_dt_strip<T, FAST>), dg, sb, 0, st, a, m0, ny, nx, dx, part);
        if (int e = launch_status()) return e;
        hipLaunchKernelGGL(k_dt_final<T>, dim3(1), dim3(FIN_THREADS), 0, st, part, int(dg.x * dg.y),
                           static_cast<T>(cfl), 1, dtp);
        return launch_status();
    }
    const uint32_t HY = (1u << PXH_G_H) | (1u << PXH_G_HUV);
    const uint32_t BED = (1u << PXH_G_ZB) | (sd->mpm ? 0u : (1u << PXH_G_G));
    if (frame_fused(g, bc)) {
        // PXH_PACKED_STAGE (float32 hardware math): b's h, hu, hv, z_b as
        // four-float records at b.h, which holds 4 ny nx elements
        const bool pk = FAST && sizeof(T) == 4 && (g->flags & PXH_PACKED_STAGE);
        F<T> bq = b;
        if (pk) {
            bq.hu = b.h + 1;
            bq.hv = b.h + 2;
            bq.z_b = b.h + 3;
        }
        const int sq = pk ? 4 : 1;
        // stage 1: b = S(a)
        if (int e = rx_stage0_fused<T, FAST>(g, a, bq, m0, sd, bc, dtp, sd->mpm ? 0 : 1, st, pk)) return e;
        if (int e = bc_launch<T>(g, bq, bc, BED, st, sq)) return e;
        // stage 2: c = S(b), a = 0.5 a + 0.5 c, and the dt of the new a
        if (int e = stage_sweep_f<T, FAST>(g, bq, c, &a, m0, sd, dtp, 1, part, st, pk)) return e;
        if (int e = bc_launch<T>(g, c, bc, HY, st)) return e;
        if (int e = frame_sweep_f<T, FAST>(g, c, bq, c, &a, m0, sd, dtp, 1, cfl, part, st, 1, sq, 1)) return e;
        return bc_launch<T>(g, a, bc, (1u << PXH_G_ZB) | (1u << PXH_G_G), st);
    }
    hipLaunchKernelGGL((k_rx_hydro_strip<T, FAST>), hg, sb, 0, st, a, b, m0, dtp, ny, nx, dx, 1);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(g, b, bc, HY, st)) return e;
    hipLaunchKernelGGL((k_rx_bed_strip<T, FAST, 0, false>), bg, sb, 0, st, b, a.z_b, a.z_b_lo, a.G, b, b, m0, dtp,
                       ny, nx, dx, sed_k(*sd), sd->mpm ? 0 : 1, nullptr, nullptr);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(g, b, bc, BED, st)) return e;
    hipLaunchKernelGGL((k_rx_hydro_strip<T, FAST>), hg, sb, 0, st, b, c, m0, dtp, ny, nx, dx, 1);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(g, c, bc, HY, st)) return e;
    hipLaunchKernelGGL((k_rx_bed_strip<T, FAST, 1, true>), bg, sb, 0, st, c, b.z_b, b.z_b_lo, b.G, a, a, m0, dtp, ny,
                       nx, dx, sed_k(*sd), 1, part, seam_cols<T>(g, ws));
    if (int e = launch_status()) return e;
    if (int e = seam_dt<T, FAST>(g, a, m0, seam_cols<T>(g, ws), part + nb, st)) return e;
    if (int e = bc_launch<T>(g, a, bc, (1u << PXH_G_ZB) | (1u << PXH_G_G), st)) return e;
    hipLaunchKernelGGL(k_dt_final<T>, dim3(1), dim3(FIN_THREADS), 0, st, part, nb + seam_blocks(ny, nx),
                       static_cast<T>(cfl), 1, dtp);
    return launch_status();
}

// The two sweeps of a stage, one launch each.
template <typename T, bool FAST>
static int rx_hydro_sweep_t(const pxh_grid *g, const pxh_fields *in_, const pxh_fields *out_, const uint8_t *m0,
                            const void *dt, hipStream_t st) {
    F<T> in = cast_fields<T>(in_), out = cast_fields<T>(out_);
    const dim3 hg = strip_grid<SO_WIDE, SR_HYDRO>(g->ny, g->nx);
    hipLaunchKernelGGL((k_rx_hydro_strip<T, FAST>), hg, dim3(SW * SWPB), 0, st, in, out, m0, static_cast<const T *>(dt),
                       g->ny, g->nx, static_cast<T>(g->dx), 1);
    return launch_status();
}

template <typename T, bool FAST>
static int rx_bed_sweep_t(const pxh_grid *g, const pxh_fields *hyd_, const pxh_fields *bed_, const pxh_fields *out_,
                          const pxh_fields *old_, const uint8_t *m0, const pxh_sediments *sd, const void *dt,
                          void *partial, hipStream_t st) {
    F<T> hyd = cast_fields<T>(hyd_), bed = cast_fields<T>(bed_), out = cast_fields<T>(out_);
    const dim3 sg = strip_grid<SO_WIDE, SR_BED>(g->ny, g->nx), sb(SW * SWPB);
    const T *dtp = static_cast<const T *>(dt);
    const T dx = static_cast<T>(g->dx);
    if (!old_) {
        hipLaunchKernelGGL((k_rx_bed_strip<T, FAST, 0, false>), sg, sb, 0, st, hyd, bed.z_b, bed.z_b_lo, bed.G, out,
                           out, m0, dtp, g->ny, g->nx, dx, sed_k(*sd), 1, nullptr, nullptr);
    } else {
        F<T> old = cast_fields<T>(old_);
        if (partial) {
            T *xs = seam_cols<T>(g, partial);
            hipLaunchKernelGGL((k_rx_bed_strip<T, FAST, 1, true>), sg, sb, 0, st, hyd, bed.z_b, bed.z_b_lo, bed.G,
                               out, old, m0, dtp, g->ny, g->nx, dx, sed_k(*sd), 1, static_cast<T *>(partial), xs);
            if (int e = launch_status()) return e;
            return seam_dt<T, FAST>(g, out, m0, xs, static_cast<T *>(partial) + sg.x * sg.y, st);
        } else
            hipLaunchKernelGGL((k_rx_bed_strip<T, FAST, 1, false>), sg, sb, 0, st, hyd, bed.z_b, bed.z_b_lo, bed.G,
                               out, old, m0, dtp, g->ny, g->nx, dx, sed_k(*sd), 1, nullptr, nullptr);
    }
    return launch_status();
}

template <typename T>
static int rx_hydro_sweep(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                          const void *dt, hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if (!(g->flags & PXH_EXACT_F32)) return rx_hydro_sweep_t<T, true>(g, in, out, m0, dt, st);
    }
    return rx_hydro_sweep_t<T, false>(g, in, out, m0, dt, st);
}

template <typename T>
static int rx_bed_sweep(const pxh_grid *g, const pxh_fields *hyd, const pxh_fields *bed, const pxh_fields *out,
                        const pxh_fields *old, const uint8_t *m0, const pxh_sediments *sd, const void *dt,
                        void *partial, hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if (!(g->flags & PXH_EXACT_F32)) return rx_bed_sweep_t<T, true>(g, hyd, bed, out, old, m0, sd, dt, partial, st);
    }
    return rx_bed_sweep_t<T, false>(g, hyd, bed, out, old, m0, sd, dt, partial, st);
}

template <typename T>
static int rx_stage_sweep(const pxh_grid *g, const pxh_fields *in_, const pxh_fields *out_, const pxh_fields *old_,
                          const uint8_t *m0, const pxh_sediments *sd, const void *dt, int write_G, void *partial,
                          hipStream_t st, bool packed) {
    const F<T> in = cast_fields<T>(in_), out = cast_fields<T>(out_);
    F<T> old{};
    if (old_) old = cast_fields<T>(old_);
    const T *dtp = static_cast<const T *>(dt);
    T *part = static_cast<T *>(partial);
    if constexpr (sizeof(T) == 4) {
        if (!(g->flags & PXH_EXACT_F32))
            return stage_sweep_f<T, true>(g, in, out, old_ ? &old : nullptr, m0, sd, dtp, write_G, part, st, packed);
    }
    return stage_sweep_f<T, false>(g, in, out, old_ ? &old : nullptr, m0, sd, dtp, write_G, part, st);
}

template <typename T>
static int rx_frame_sweep(const pxh_grid *g, const pxh_fields *hyd_, const pxh_fields *bed_, const pxh_fields *out_,
                          const pxh_fields *old_, const uint8_t *m0, const pxh_sediments *sd, void *dt, int write_G,
                          double cfl, void *partial, hipStream_t st, int hs, int zs, int os) {
    const F<T> hyd = cast_fields<T>(hyd_), bed = cast_fields<T>(bed_);
    F<T> out{}, old{};
    if (out_) out = cast_fields<T>(out_);
    if (old_) old = cast_fields<T>(old_);
    T *dtp = static_cast<T *>(dt);
    T *part = static_cast<T *>(partial);
    if constexpr (sizeof(T) == 4) {
        if (!(g->flags & PXH_EXACT_F32))
            return frame_sweep_f<T, true>(g, hyd, bed, out, old_ ? &old : nullptr, m0, sd, dtp, write_G, cfl, part, st,
                                          hs, zs, os);
    }
    return frame_sweep_f<T, false>(g, hyd, bed, out, old_ ? &old : nullptr, m0, sd, dtp, write_G, cfl, part, st, hs,
                                   zs, os);
}

template <typename T>
static int rx_step(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                   const pxh_sediments *sd, const pxh_bc *bc, const void *dt, hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if (!(g->flags & PXH_EXACT_F32)) return rx_step_t<T, true>(g, in, out, m0, sd, bc, dt, st);
    }
    return rx_step_t<T, false>(g, in, out, m0, sd, bc, dt, st);
}

template <typename T>
static int rx_ssprk2(const pxh_grid *g, const pxh_fields *a, const pxh_fields *b, const pxh_fields *c,
                     const uint8_t *m0, const pxh_sediments *sd, const pxh_bc *bc, double cfl, void *dt, void *ws,
                     int dt_ready, hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if (!(g->flags & PXH_EXACT_F32))
            return rx_ssprk2_t<T, true>(g, a, b, c, m0, sd, bc, cfl, dt, ws, dt_ready, st);
    }
    return rx_ssprk2_t<T, false>(g, a, b, c, m0, sd, bc, cfl, dt, ws, dt_ready, st);
}

template <typename T, bool FAST>
static int rx_dt_t(const pxh_grid *g, const pxh_fields *s_, const uint8_t *m0, double cfl, void *dt_out, void *ws,
                   hipStream_t st) {
    F<T> s = cast_fields<T>(s_);
    const dim3 sg = strip_grid<SO_WIDE, SR_BED>(g->ny, g->nx);
    const int nb = int(sg.x * sg.y);
    T *part = static_cast<T *>(ws);
    hipLaunchKernelGGL((k_rx_dt_strip<T, FAST>), sg, dim3(SW * SWPB), 0, st, s, m0, g->ny, g->nx,
                       static_cast<T>(g->dx), part);
    if (int e = launch_status()) return e;
    hipLaunchKernelGGL(k_dt_final<T>, dim3(1), dim3(FIN_THREADS), 0, st, part, nb, static_cast<T>(cfl), 1,
                       static_cast<T *>(dt_out));
    return launch_status();
}

template <typename T>
static int rx_dt(const pxh_grid *g, const pxh_fields *s, const uint8_t *m0, double cfl, void *dt_out, void *ws,
                 hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if (!(g->flags & PXH_EXACT_F32)) return rx_dt_t<T, true>(g, s, m0, cfl, dt_out, ws, st);
    }
    return rx_dt_t<T, false>(g, s, m0, cfl, dt_out, ws, st);
}

template <typename T>
static int roe_dt(const pxh_grid *g, const pxh_fields *s_, double cfl, void *dt_out, void *ws, hipStream_t st) {
    F<T> s = cast_fields<T>(s_);
    const int nb = dt_blocks(g);
    T *part = static_cast<T *>(ws);
    hipLaunchKernelGGL(k_roe_dt_partial<T>, dim3(nb), dim3(DT_THREADS), 0, st, s, long(g->ny) * g->nx,
                       static_cast<T>(g->dx), part);
    if (int e = launch_status()) return e;
    hipLaunchKernelGGL(k_dt_final<T>, dim3(1), dim3(FIN_THREADS), 0, st, part, nb, static_cast<T>(cfl), 0,
                       static_cast<T *>(dt_out));
    return launch_status();
}

template <typename T>
static int roe_step(const pxh_grid *g, const pxh_fields *in_, const pxh_fields *out_, const pxh_bc *bc,
                    const void *dt, hipStream_t st) {
    F<T> in = cast_fields<T>(in_), out = cast_fields<T>(out_);
    hipLaunchKernelGGL(k_roe_hydro<T>, grid2d(g->ny, g->nx), dim3(BX, BY), 0, st, in, out,
                       static_cast<const T *>(dt), g->ny, g->nx, static_cast<T>(g->dx));
    if (int e = launch_status()) return e;
    return bc_launch<T>(g, out, bc, (1u << PXH_G_H) | (1u << PXH_G_HUV), st);
}

template <typename T>
static int bc_apply_t(const pxh_grid *g, const pxh_fields *s_, const pxh_bc *bc, hipStream_t st, int sq) {
    return bc_launch<T>(g, cast_fields<T>(s_), bc, 0xffffffffu, st, sq);
}

// A pxh_fields whose hu, hv, z_b point one, two and three elements past h is
// a view of four-element records (h, hu, hv, z_b of a cell side by side): the
// layout the float32 fused step keeps its first stage's output in
// (PXH_PACKED_STAGE), which the separate sweep entries and the ring fill
// accept where that step makes it.
static bool packed_view(const pxh_grid *g, const pxh_fields *f) {
    if (!f || !f->h) return false;
    const long es = g->dtype == PXH_F64 ? 8 : 4;
    const char *h = static_cast<const char *>(f->h);
    return f->hu == h + es && f->hv == h + 2 * es && f->z_b == h + 3 * es;
}

// A compensated bed is float32-only and all-or-nothing across the buffers of a call.
static bool lo_ok(const pxh_grid *g, std::initializer_list<const pxh_fields *> fs) {
    int n = 0, k = 0;
    for (const pxh_fields *f : fs) {
        ++k;
        if (f->z_b_lo) ++n;
    }
    return n == 0 || (n == k && g->dtype == PXH_F32);
}

template <typename T>
static int rx_part(int which, const pxh_grid *g, const pxh_fields *in_, const pxh_fields *out_, const uint8_t *m0,
                   const pxh_sediments *sd, const void *dt, hipStream_t st) {
    F<T> in = cast_fields<T>(in_), out = cast_fields<T>(out_);
    const int ny = g->ny, nx = g->nx, ncell = ny * nx;
    const T dx = static_cast<T>(g->dx);
    const T *dtp = static_cast<const T *>(dt);
    switch (which) {
        case 0: hipLaunchKernelGGL(k_rx_hydro<T>, grid2d(ny, nx), dim3(BX, BY), 0, st, in, out, m0, dtp, ny, nx, dx); break;
        case 1: hipLaunchKernelGGL(k_rx_momcorr<T>, grid2d(ny, nx), dim3(BX, BY), 0, st, in, out, m0, ny, nx); break;
        case 2: hipLaunchKernelGGL(k_rx_G<T>, dim3((ncell + 255) / 256), dim3(256), 0, st, in, out, m0, ncell, *sd); break;
        case 3: hipLaunchKernelGGL(k_rx_exner<T>, grid2d(ny, nx), dim3(BX, BY), 0, st, in, out, m0, dtp, ny, nx, dx); break;
        default: return PXH_E_ARG;
    }
    return launch_status();
}

static int part_args_ok(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0) {
    return grid_ok(g) && fields_ok(in, true) && fields_ok(out, true) && m0;
}

// The halo plan of a call: the selected fields (PXH_G_* order, z_b_lo after
// z_b) and sides; false when the arguments do not describe one.
template <typename T>
static bool halo_plan(const pxh_grid *g, const pxh_fields *s_, int hw, uint32_t groups, int sides, HaloPlan &plan,
                      HaloFields<T> &hf) {
    if (hw < 1 || hw > 3 || g->ny < 2 * hw + 1 || g->nx < 2 * hw + 1) return false;
    // one phase per call: west/east, or north/south (their strips overlap in the corners)
    const bool we = sides & (PXH_HALO_W | PXH_HALO_E), ns = sides & (PXH_HALO_N | PXH_HALO_S);
    if (sides <= 0 || sides > 15 || (we && ns) || groups == 0 || (groups & ~0x7Fu)) return false;
    const F<T> f = cast_fields<T>(s_);
    T *list[HALO_MAX_FIELDS];
    int n = 0;
    auto add = [&](T *p) { list[n++] = p; };
    if (groups & (1u << PXH_G_H)) add(f.h);
    if (groups & (1u << PXH_G_HUV)) { add(f.hu); add(f.hv); }
    if (groups & (1u << PXH_G_Z)) add(f.z);
    if (groups & (1u << PXH_G_ZB)) { add(f.z_b); if (f.z_b_lo) add(f.z_b_lo); }
    if (groups & (1u << PXH_G_N)) add(f.n);
    if (groups & (1u << PXH_G_NB)) add(f.n_b);
    if (groups & (1u << PXH_G_G)) add(f.G);
    for (int k = 0; k < n; ++k)
        if (!list[k]) return false;
    plan.nf = n;
    long so = 0;
    for (int sd = 0; sd < 4; ++sd) {
        plan.side_on[sd] = (sides >> sd) & 1;
        plan.side_off[sd] = so;
        if (plan.side_on[sd]) so += long(hw) * ((sd == 0 || sd == 2) ? g->ny : g->nx);
    }
    plan.side_off[4] = so;
    for (int k = 0; k <= n; ++k) plan.off[k] = long(k) * so;
    for (int k = 0; k < HALO_MAX_FIELDS; ++k) hf.p[k] = k < n ? list[k] : nullptr;
    return true;
}

template <typename T>
static int halo_t(bool unpack, const pxh_grid *g, const pxh_fields *s, int hw, uint32_t groups, int sides, void *buf,
                  hipStream_t st) {
    HaloPlan plan;
    HaloFields<T> hf;
    if (!halo_plan<T>(g, s, hw, groups, sides, plan, hf)) return PXH_E_ARG;
    const long total = plan.off[plan.nf];
    const unsigned nb = unsigned(total / 256 + 1 < 4096 ? total / 256 + 1 : 4096);
    if (unpack)
        hipLaunchKernelGGL((k_halo<T, true>), dim3(nb), dim3(256), 0, st, hf, plan, static_cast<T *>(buf), g->ny,
                           g->nx, hw);
    else
        hipLaunchKernelGGL((k_halo<T, false>), dim3(nb), dim3(256), 0, st, hf, plan, static_cast<T *>(buf), g->ny,
                           g->nx, hw);
    return launch_status();
}

// ---------------------------------------------------------------------------
// ensembles (ensemble.hpp): the plan blob and the batched step
// ---------------------------------------------------------------------------
struct EnsPlanHeader {
    uint32_t magic;
    int32_t n, dtype, flags;
    int32_t mpm;           // every member's sediment mode
    int32_t sblocks;       // stage-sweep blocks over all members
    int32_t fblocks;       // frame blocks
    int32_t dblocks;       // standalone-dt blocks
    int32_t nparts;        // partial minima
    int32_t tblocks;       // frame + seam dt blocks
    int64_t tab_off, sed_off, bytes;
    int64_t cells;         // arena elements the members span
    int64_t xs_elems;      // seam-column buffers of all members
};
constexpr uint32_t ENS_MAGIC = 0x50584845u;

static int64_t align16(int64_t v) { return (v + 15) & ~int64_t(15); }

static int64_t ens_plan_build(int32_t dtype, int32_t flags, const pxh_member *mem, int32_t n, void *out,
                              int64_t out_bytes) {
    if (n <= 0 || !mem || (dtype != PXH_F32 && dtype != PXH_F64)) return PXH_E_ARG;
    const int rows = PXH_STRIP_ROWS(flags);
    if (rows != 0 && (rows < 8 || rows > 128)) return PXH_E_ARG;
    const int64_t tab_off = align16(sizeof(EnsPlanHeader));
    const int64_t sed_off = align16(tab_off + int64_t(n + 1) * sizeof(EnsMember));
    const int64_t bytes = align16(sed_off + int64_t(n) * sizeof(SedK));
    if (!out || out_bytes < bytes) return bytes;
    char *base = static_cast<char *>(out);
    EnsPlanHeader h{};
    h.magic = ENS_MAGIC;
    h.n = n;
    h.dtype = dtype;
    h.flags = flags;
    h.mpm = mem[0].sed.mpm;
    EnsMember *tab = reinterpret_cast<EnsMember *>(base + tab_off);
    SedK *seds = reinterpret_cast<SedK *>(base + sed_off);
    int64_t sb = 0, fb = 0, db = 0, tb = 0, np = 0, cells = 0, xe = 0;
    const long esz = dtype == PXH_F64 ? 8 : 4;
    for (int i = 0; i < n; ++i) {
        const pxh_member &q = mem[i];
        if (q.ny < 2 || q.nx < 2 || q.off < 0 || !(q.dx > 0)) return PXH_E_ARG;
        if (long(strip_rows(flags) + 4) * q.nx * esz >= (1L << 32)) return PXH_E_RANGE;
        const pxh_sediments &sd = q.sed;
        if (sd.n_sed < 0 || sd.n_sed > PXH_MAX_SED || (sd.mpm && sd.n_sed < 1) || sd.mpm != h.mpm) return PXH_E_RANGE;
        EnsMember m{};
        m.off = q.off;
        m.ny = q.ny;
        m.nx = q.nx;
        m.dx = q.dx;
        const dim3 fg = stage_grid(q.ny, q.nx, strip_rows(flags)), dg = strip_grid<SO_WIDE, SR_BED>(q.ny, q.nx);
        m.sgx = int(fg.x);
        m.nstage = int(fg.x * fg.y);
        m.nband = frame_cells(q.ny, q.nx);
        m.nframe = int((m.nband + 255) / 256);
        m.ntdt = int((band_dt_work(q.ny, q.nx, strip_rows(flags)) + 255) / 256);
        m.dgx = int(dg.x);
        m.ndt = int(dg.x * dg.y);
        m.sblk0 = int(sb);
        m.fblk0 = int(fb);
        m.dblk0 = int(db);
        m.tblk0 = int(tb);
        m.part0 = int(np);
        m.xs0 = xe;
        xe += seam_col_elems(q.ny, q.nx);
        sb += m.nstage;
        fb += m.nframe;
        db += m.ndt;
        tb += m.ntdt;
        np += (m.nstage + m.ntdt > m.ndt ? m.nstage + m.ntdt : m.ndt);
        const int64_t end = q.off + int64_t(q.ny) * q.nx;
        if (end > cells) cells = end;
        tab[i] = m;
        seds[i] = sed_k(sd);
    }
    if (sb >= (1L << 31) || tb >= (1L << 31) || np >= (1L << 31) || cells >= (1L << 31)) return PXH_E_RANGE;
    EnsMember sentinel{};
    sentinel.sblk0 = int(sb);
    sentinel.fblk0 = int(fb);
    sentinel.dblk0 = int(db);
    sentinel.tblk0 = int(tb);
    sentinel.part0 = int(np);
    tab[n] = sentinel;
    h.sblocks = int(sb);
    h.fblocks = int(fb);
    h.dblocks = int(db);
    h.tblocks = int(tb);
    h.nparts = int(np);
    h.tab_off = tab_off;
    h.sed_off = sed_off;
    h.bytes = bytes;
    h.cells = cells;
    h.xs_elems = xe;
    *reinterpret_cast<EnsPlanHeader *>(base) = h;
    return bytes;
}

// The ensemble workspace: partial minima (8 bytes each), then the members'
// seam-column buffers from this byte offset.
static int64_t ens_xs_offset(const EnsPlanHeader &h) { return (int64_t(h.nparts) * 8 + 255) & ~int64_t(255); }

static const EnsPlanHeader *ens_header(const void *plan) {
    const EnsPlanHeader *h = static_cast<const EnsPlanHeader *>(plan);
    return (h && h->magic == ENS_MAGIC && h->n > 0) ? h : nullptr;
}

template <typename T, bool FAST, int MODE, bool DT>
static void ens_stage(const EnsPlanHeader &h, const EnsMember *tab, const SedK *seds, hipStream_t st, const F<T> &in,
                      const F<T> &out, const F<T> &cb, const F<T> &old, const uint8_t *m0, const T *dts,
                      int write_G, T *partial, T *xs, bool pk = false) {
    const bool lo = sizeof(T) == 4 && in.z_b_lo != nullptr;
    constexpr bool has_pk = FAST && sizeof(T) == 4;
    // the plan's PXH_NO_NODATA / PXH_Z_UNIFORM promises, as launch_stage takes the grid's
    const bool nn = (h.flags & PXH_NO_NODATA) != 0, zu = nn && (h.flags & PXH_Z_UNIFORM) != 0;
    constexpr bool has_zu = FAST || sizeof(T) == 8;
#define PXH_ENS_PK(LO_, MPM_, ST_, PK_)                                                                         \
    hipLaunchKernelGGL((k_ens_stage<T, FAST, MODE, DT, LO_, MPM_, ST_, PK_>), dim3(h.sblocks), dim3(SW * FWPB), 0, \
                       st, in, out, cb, old, m0, dts, tab, h.n, strip_rows(h.flags), seds, write_G, partial, xs)
#define PXH_ENS_ST(LO_, MPM_, ST_)                   \
    do {                                             \
        if constexpr (has_pk) {                      \
            if (pk) {                                \
                PXH_ENS_PK(LO_, MPM_, ST_, true);    \
                break;                               \
            }                                        \
        }                                            \
        PXH_ENS_PK(LO_, MPM_, ST_, false);           \
    } while (0)
#define PXH_ENS(LO_, MPM_)                              \
    do {                                                \
        if constexpr (has_zu) {                         \
            if (zu) {                                   \
                PXH_ENS_ST(LO_, MPM_, ST_NN | ST_ZU);   \
                break;                                  \
            }                                           \
        }                                               \
        if (nn) PXH_ENS_ST(LO_, MPM_, ST_NN);           \
        else PXH_ENS_ST(LO_, MPM_, 0);                  \
    } while (0)
    if constexpr (sizeof(T) == 4) {
        if (lo) {
            if (h.mpm) PXH_ENS(true, true);
            else PXH_ENS(true, false);
            return;
        }
    }
    if (h.mpm) PXH_ENS(false, true);
    else PXH_ENS(false, false);
#undef PXH_ENS
#undef PXH_ENS_ST
#undef PXH_ENS_PK
}

template <typename T, bool FAST>
static int ens_dt_t(const EnsPlanHeader &h, const char *pd, const F<T> &a, const uint8_t *m0, double cfl, T *dts,
                    T *part, hipStream_t st) {
    const EnsMember *tab = reinterpret_cast<const EnsMember *>(pd + h.tab_off);
    hipLaunchKernelGGL((k_ens_dt_strip<T, FAST>), dim3(h.dblocks), dim3(SW * SWPB), 0, st, a, m0, tab, h.n, part);
    if (int e = launch_status()) return e;
    hipLaunchKernelGGL(k_ens_dt_final<T>, dim3(h.n), dim3(256), 0, st, part, tab, static_cast<T>(cfl), 0, dts);
    return launch_status();
}

template <typename T, bool FAST>
static int ens_ssprk2_t(const EnsPlanHeader &h, const char *pd, const pxh_fields *a_, const pxh_fields *b_,
                        const pxh_fields *c_, const uint8_t *m0, const pxh_bc *bc, double cfl, T *dts, T *part,
                        int dt_ready, hipStream_t st) {
    const F<T> a = cast_fields<T>(a_), b = cast_fields<T>(b_), c = cast_fields<T>(c_);
    const EnsMember *tab = reinterpret_cast<const EnsMember *>(pd + h.tab_off);
    const SedK *seds = reinterpret_cast<const SedK *>(pd + h.sed_off);
    const F<T> none{};
    const pxh_grid g{1, 1, h.dtype, h.flags, 1.0};  // the ring fills read no geometry
    if (!dt_ready)
        if (int e = ens_dt_t<T, FAST>(h, pd, a, m0, cfl, dts, part, st)) return e;
    const uint32_t HY = (1u << PXH_G_H) | (1u << PXH_G_HUV);
    const uint32_t BED = (1u << PXH_G_ZB) | (h.mpm ? 0u : (1u << PXH_G_G));
    // PXH_PACKED_STAGE (float32 hardware math): b's h, hu, hv, z_b as
    // four-float records over the arena at b.h, which holds 4 cells' worth
    const bool pk = FAST && sizeof(T) == 4 && (h.flags & PXH_PACKED_STAGE);
    F<T> bq = b;
    if (pk) {
        bq.hu = b.h + 1;
        bq.hv = b.h + 2;
        bq.z_b = b.h + 3;
    }
    const int sq = pk ? 4 : 1;
    // stage 1: b = S(a)
    T *xs = reinterpret_cast<T *>(reinterpret_cast<char *>(part) + ens_xs_offset(h));
    ens_stage<T, FAST, 0, false>(h, tab, seds, st, a, bq, none, none, m0, dts, h.mpm ? 0 : 1, nullptr, nullptr, pk);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(&g, bq, bc, HY, st, sq)) return e;
    hipLaunchKernelGGL((k_ens_band_bed<T, FAST, 0>), dim3(h.fblocks), dim3(256), 0, st, bq, a, bq, none, m0, dts,
                       tab, h.n, seds, h.mpm ? 0 : 1, sq, 1, sq);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(&g, bq, bc, BED, st, sq)) return e;
    // stage 2: c = S(b), a = 0.5 a + 0.5 c, and each member's dt of the new a
    ens_stage<T, FAST, 1, true>(h, tab, seds, st, bq, none, c, a, m0, dts, 1, part, xs, pk);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(&g, c, bc, HY, st)) return e;
    hipLaunchKernelGGL((k_ens_band_bed<T, FAST, 1>), dim3(h.fblocks), dim3(256), 0, st, c, bq, none, a, m0, dts,
                       tab, h.n, seds, 1, 1, sq, 1);
    if (int e = launch_status()) return e;
    hipLaunchKernelGGL((k_ens_band_dt<T, FAST>), dim3(h.tblocks), dim3(256), 0, st, a, m0, tab, h.n,
                       strip_rows(h.flags), xs, part);
    if (int e = launch_status()) return e;
    if (int e = bc_launch<T>(&g, a, bc, (1u << PXH_G_ZB) | (1u << PXH_G_G), st)) return e;
    hipLaunchKernelGGL(k_ens_dt_final<T>, dim3(h.n), dim3(256), 0, st, part, tab, static_cast<T>(cfl), 1, dts);
    return launch_status();
}

template <typename T>
static int ens_ssprk2(const EnsPlanHeader &h, const char *pd, const pxh_fields *a, const pxh_fields *b,
                      const pxh_fields *c, const uint8_t *m0, const pxh_bc *bc, double cfl, void *dts, void *ws,
                      int dt_ready, hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if (!(h.flags & PXH_EXACT_F32))
            return ens_ssprk2_t<T, true>(h, pd, a, b, c, m0, bc, cfl, static_cast<T *>(dts), static_cast<T *>(ws),
                                         dt_ready, st);
    }
    return ens_ssprk2_t<T, false>(h, pd, a, b, c, m0, bc, cfl, static_cast<T *>(dts), static_cast<T *>(ws),
                                  dt_ready, st);
}

template <typename T>
static int ens_dt(const EnsPlanHeader &h, const char *pd, const pxh_fields *a, const uint8_t *m0, double cfl,
                  void *dts, void *ws, hipStream_t st) {
    const F<T> f = cast_fields<T>(a);
    if constexpr (sizeof(T) == 4) {
        if (!(h.flags & PXH_EXACT_F32))
            return ens_dt_t<T, true>(h, pd, f, m0, cfl, static_cast<T *>(dts), static_cast<T *>(ws), st);
    }
    return ens_dt_t<T, false>(h, pd, f, m0, cfl, static_cast<T *>(dts), static_cast<T *>(ws), st);
}

}  // namespace pxh

using namespace pxh;

#define DISPATCH(g, fn, ...) ((g)->dtype == PXH_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

// The multi-domain fused sweeps (pxh_roe_exner_stage_sweep_multi / _frame_sweep_multi).
// flags: PXH_NO_NODATA / PXH_Z_UNIFORM, promises for every domain, picked as
// launch_stage picks them for one grid.
template <typename T, bool FAST, int MODE, bool LO, bool MPM, bool DT>
static void multi_stage_k(const pxh_sweep_domain *doms, int n, int nblocks, int sr, int flags, const T *dtp,
                          const SedK &sd, int write_G, T *partial, T *xs, hipStream_t st) {
    const bool nn = (flags & PXH_NO_NODATA) != 0, zu = nn && (flags & PXH_Z_UNIFORM) != 0;
    const dim3 grid(nblocks), block(SW * FWPB);
    if constexpr (FAST || sizeof(T) == 8) {
        if (zu) {
            hipLaunchKernelGGL((k_multi_stage<T, FAST, MODE, LO, MPM, ST_NN | ST_ZU, DT>), grid, block, 0, st, doms,
                               n, sr, dtp, sd, write_G, partial, xs);
            return;
        }
    }
    if (nn)
        hipLaunchKernelGGL((k_multi_stage<T, FAST, MODE, LO, MPM, ST_NN, DT>), grid, block, 0, st, doms, n, sr, dtp,
                           sd, write_G, partial, xs);
    else
        hipLaunchKernelGGL((k_multi_stage<T, FAST, MODE, LO, MPM, 0, DT>), grid, block, 0, st, doms, n, sr, dtp, sd,
                           write_G, partial, xs);
}

// partial, xs: the second stage's dt minima and seam records (both or neither).
template <typename T, bool FAST>
static void multi_stage_t(const pxh_sweep_domain *doms, int n, int nblocks, int sr, bool second, bool lo, int flags,
                          const T *dtp, const SedK &sd, int write_G, T *partial, T *xs, hipStream_t st) {
#define PXH_MS(MODE_, LO_, MPM_, DT_) \
    multi_stage_k<T, FAST, MODE_, LO_, MPM_, DT_>(doms, n, nblocks, sr, flags, dtp, sd, write_G, partial, xs, st)
#define PXH_MS_MODE(LO_, MPM_)                   \
    if (second && partial) PXH_MS(1, LO_, MPM_, true); \
    else if (second) PXH_MS(1, LO_, MPM_, false); \
    else PXH_MS(0, LO_, MPM_, false)
    if constexpr (sizeof(T) == 4) {
        if (lo) {
            if (sd.mpm) { PXH_MS_MODE(true, true); }
            else { PXH_MS_MODE(true, false); }
            return;
        }
    }
    if (sd.mpm) { PXH_MS_MODE(false, true); }
    else { PXH_MS_MODE(false, false); }
#undef PXH_MS_MODE
#undef PXH_MS
}

// The frame and seam edges of every domain (k_multi_band_dt, which folds in
// the second stage's per-block minima in partial[0, nstage)), then the final
// reduction of its own partials.
template <typename T, bool FAST>
static int dt_frame_multi_t(const pxh_sweep_domain *doms, int n, int nblocks, int nstage, int sr, T cfl,
                            T *partial, const T *xs, T *dt_out, hipStream_t st) {
    hipLaunchKernelGGL((k_multi_band_dt<T, FAST>), dim3(nblocks), dim3(256), 0, st, doms, n, sr, xs,
                       static_cast<const T *>(partial), nstage, partial + nstage);
    if (int e = launch_status()) return e;
    hipLaunchKernelGGL(k_dt_final<T>, dim3(1), dim3(FIN_THREADS), 0, st, static_cast<const T *>(partial + nstage),
                       nblocks, cfl, 1, dt_out);
    return launch_status();
}

template <typename T, bool FAST>
static void multi_frame_t(const pxh_sweep_domain *doms, int n, int nblocks, bool second, const T *dtp,
                          const SedK &sd, int write_G, hipStream_t st) {
    if (second)
        hipLaunchKernelGGL((k_multi_band_bed<T, FAST, 1>), dim3(nblocks), dim3(256), 0, st, doms, n, dtp, sd, write_G);
    else
        hipLaunchKernelGGL((k_multi_band_bed<T, FAST, 0>), dim3(nblocks), dim3(256), 0, st, doms, n, dtp, sd, write_G);
}

// ---------------------------------------------------------------------------
// The dt workspace of pxh_roe_exner_ssprk2_noalloc: a static device array,
// one copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (in host code it names a host-side stand-in): it
// passes the device address hipGetSymbolAddress returns, looked up once per
// device with that device current and kept in scratch_dev[].  Every write of
// the workspace precedes its reads within one call (each block stores its
// partial minimum and seam records before the reductions read them), so it
// needs no clearing between calls.  Calls that overlap on one GPU share it
// and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PXH_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static void *scratch_dev[SCRATCH_MAX_DEVICES];
static std::mutex scratch_mu;

// The device address of `scratch` on the device that owns `st` (the current
// device for the null stream).
static int scratch_of(hipStream_t st, void **out) {
    int dev = 0;
    hipError_t e = st ? hipStreamGetDevice(st, &dev) : hipGetDevice(&dev);
    if (e != hipSuccess) return int(e);
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return PXH_E_RANGE;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        int cur = 0;
        if ((e = hipGetDevice(&cur)) != hipSuccess) return int(e);
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return int(e);
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) {
            const hipError_t e2 = hipSetDevice(cur);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return int(e);
        if (!p) return PXH_E_ARG;
        scratch_dev[dev] = p;
    }
    *out = scratch_dev[dev];
    return PXH_OK;
}

extern "C" {

int pxh_version(void) { return 1; }

const char *pxh_error_string(int status) {
    if (status == PXH_OK) return "ok";
    if (status == PXH_E_ARG) return "invalid argument";
    if (status == PXH_E_RANGE) return "size beyond a compiled-in limit";
    if (status == PXH_E_UNSUPPORTED) return "unsupported configuration";
    if (status > 0) return hipGetErrorString(static_cast<hipError_t>(status));
    return "unknown error";
}

int pxh_roe_exner_init(const pxh_grid *g, const pxh_fields *s, const uint8_t *m0, const pxh_sediments *sd,
                       void *stream) {
    if (!grid_ok(g) || !fields_ok(s, true) || !m0 || !sd) return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    const long n = long(g->ny) * g->nx;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (g->dtype == PXH_F64)
        hipLaunchKernelGGL(k_rx_init<double>, dim3((n + 255) / 256), dim3(256), 0, st, cast_fields<double>(s), m0, n,
                           *sd);
    else
        hipLaunchKernelGGL(k_rx_init<float>, dim3((n + 255) / 256), dim3(256), 0, st, cast_fields<float>(s), m0, n,
                           *sd);
    return launch_status();
}

int pxh_roe_exner_step(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                       const pxh_sediments *sd, const pxh_bc *bc, const void *dt_dev, void *stream) {
    if (!grid_ok(g) || !fields_ok(in, true) || !fields_ok(out, true) || !m0 || !sd || !dt_dev) return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    if (in->h == out->h || in->hu == out->hu || in->hv == out->hv || in->z_b == out->z_b || in->G == out->G)
        return PXH_E_ARG;
    if (!lo_ok(g, {in, out}) || (in->z_b_lo && in->z_b_lo == out->z_b_lo)) return PXH_E_ARG;
    if (!bc_ok(bc)) return PXH_E_ARG;
    if (bc && bc->n_levels > 1) return PXH_E_UNSUPPORTED;
    return DISPATCH(g, rx_step, g, in, out, m0, sd, bc, dt_dev, static_cast<hipStream_t>(stream));
}

int pxh_roe_solve_2D(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                     const void *dt_dev, void *stream) {
    if (!part_args_ok(g, in, out, m0) || !dt_dev || in->h == out->h || in->hu == out->hu || in->hv == out->hv)
        return PXH_E_ARG;
    return DISPATCH(g, rx_part, 0, g, in, out, m0, nullptr, dt_dev, static_cast<hipStream_t>(stream));
}

int pxh_momentum_corrections(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                             void *stream) {
    if (!part_args_ok(g, in, out, m0)) return PXH_E_ARG;
    return DISPATCH(g, rx_part, 1, g, in, out, m0, nullptr, nullptr, static_cast<hipStream_t>(stream));
}

int pxh_compute_G(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                  const pxh_sediments *sd, void *stream) {
    if (!part_args_ok(g, in, out, m0) || !sd) return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    return DISPATCH(g, rx_part, 2, g, in, out, m0, sd, nullptr, static_cast<hipStream_t>(stream));
}

int pxh_exner_solve_2D(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                       const void *dt_dev, void *stream) {
    if (!part_args_ok(g, in, out, m0) || !dt_dev || in->z_b == out->z_b) return PXH_E_ARG;
    return DISPATCH(g, rx_part, 3, g, in, out, m0, nullptr, dt_dev, static_cast<hipStream_t>(stream));
}

int pxh_roe_exner_hydro_sweep(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                              const void *dt_dev, void *stream) {
    if (!part_args_ok(g, in, out, m0) || !dt_dev || in->h == out->h || in->hu == out->hu || in->hv == out->hv)
        return PXH_E_ARG;
    return DISPATCH(g, rx_hydro_sweep, g, in, out, m0, dt_dev, static_cast<hipStream_t>(stream));
}

int pxh_roe_exner_bed_sweep(const pxh_grid *g, const pxh_fields *hyd, const pxh_fields *bed, const pxh_fields *out,
                            const pxh_fields *old, const uint8_t *m0, const pxh_sediments *sd, const void *dt_dev,
                            void *dt_partial, void *stream) {
    if (!grid_ok(g) || !fields_ok(hyd, true) || !fields_ok(bed, true) || !fields_ok(out, true) || !m0 || !sd ||
        !dt_dev)
        return PXH_E_ARG;
    if (old && !fields_ok(old, true)) return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    if (bed->z_b == out->z_b || (bed->z_b_lo && bed->z_b_lo == out->z_b_lo)) return PXH_E_ARG;
    if (!lo_ok(g, {bed, out}) || (old && !lo_ok(g, {bed, out, old}))) return PXH_E_ARG;
    if (old && (hyd->h == out->h || hyd->hu == out->hu || hyd->hv == out->hv)) return PXH_E_ARG;
    return DISPATCH(g, rx_bed_sweep, g, hyd, bed, out, old, m0, sd, dt_dev, dt_partial,
                    static_cast<hipStream_t>(stream));
}

int pxh_count_nonfinite(const pxh_grid *g, const pxh_fields *s, void *count_dev, void *stream) {
    if (!grid_ok(g) || !s || !s->h || !s->hu || !s->hv || !count_dev) return PXH_E_ARG;
    if (s->z_b_lo && !lo_ok(g, {s})) return PXH_E_ARG;
    auto *cnt = static_cast<unsigned long long *>(count_dev);
    hipStream_t st = static_cast<hipStream_t>(stream);
    return g->dtype == PXH_F64 ? count_nonfinite_t<double>(g, s, cnt, st) : count_nonfinite_t<float>(g, s, cnt, st);
}

int pxh_dt_finalize(const pxh_grid *g, const void *dt_partial, double cfl, void *dt_out, void *stream) {
    if (!grid_ok(g) || !dt_partial || !dt_out) return PXH_E_ARG;
    const dim3 sg = strip_grid<SO_WIDE, SR_BED>(g->ny, g->nx);  // the partials of pxh_roe_exner_bed_sweep
    const int nb = int(sg.x * sg.y) + seam_blocks(g->ny, g->nx);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (g->dtype == PXH_F64)
        hipLaunchKernelGGL(k_dt_final<double>, dim3(1), dim3(FIN_THREADS), 0, st, static_cast<const double *>(dt_partial),
                           nb, cfl, 1, static_cast<double *>(dt_out));
    else
        hipLaunchKernelGGL(k_dt_final<float>, dim3(1), dim3(FIN_THREADS), 0, st, static_cast<const float *>(dt_partial),
                           nb, static_cast<float>(cfl), 1, static_cast<float *>(dt_out));
    return launch_status();
}

int pxh_roe_exner_stage_sweep(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const pxh_fields *old,
                              const uint8_t *m0, const pxh_sediments *sd, const void *dt_dev, int flags,
                              void *dt_partial, void *stream) {
    if (!grid_ok(g) || !fields_ok(in, true, true) || !fields_ok(out, false, true) || !m0 || !sd || !dt_dev)
        return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    if (in->h == out->h || in->hu == out->hu || in->hv == out->hv) return PXH_E_ARG;
    if (!old) {
        if (!fields_ok(out, true, true) || in->z_b == out->z_b || in->G == out->G || !lo_ok(g, {in, out}) ||
            (in->z_b_lo && in->z_b_lo == out->z_b_lo))
            return PXH_E_ARG;
    } else {
        if (!fields_ok(old, true) || !lo_ok(g, {in, old})) return PXH_E_ARG;
        const void *po[] = {old->h, old->hu, old->hv, old->z_b, old->G};
        for (const void *q : po)
            if (q == in->h || q == in->hu || q == in->hv || q == in->z_b || q == in->G || q == out->h ||
                q == out->hu || q == out->hv)
                return PXH_E_ARG;
        if (old->z_b_lo && old->z_b_lo == in->z_b_lo) return PXH_E_ARG;
    }
    // records: the first stage's output or the second stage's input, on the
    // float32 hardware-math path
    const bool pin = packed_view(g, in), pout = packed_view(g, out);
    if ((pin || pout) && (g->dtype != PXH_F32 || (g->flags & PXH_EXACT_F32) || (pin && !old) || (pout && old) ||
                          (pin && pout)))
        return PXH_E_ARG;
    if (packed_view(g, old)) return PXH_E_ARG;
    if (!stage_fits(g)) return PXH_E_RANGE;
    const int write_G = (flags & PXH_SWEEP_SKIP_G) && sd->mpm ? 0 : 1;
    return DISPATCH(g, rx_stage_sweep, g, in, out, old, m0, sd, dt_dev, write_G, old ? dt_partial : nullptr,
                    static_cast<hipStream_t>(stream), pin || pout);
}

int pxh_roe_exner_frame_sweep(const pxh_grid *g, const pxh_fields *hyd, const pxh_fields *bed, const pxh_fields *out,
                              const pxh_fields *old, const uint8_t *m0, const pxh_sediments *sd, void *dt_dev,
                              int flags, double cfl, void *dt_partial, void *stream) {
    if (!grid_ok(g) || !fields_ok(hyd, false, true) || !fields_ok(bed, true, true) || !m0 || !sd || !dt_dev)
        return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    if (!old) {
        if (!fields_ok(out, true, true) || bed->z_b == out->z_b || bed->G == out->G || !lo_ok(g, {bed, out}) ||
            (bed->z_b_lo && bed->z_b_lo == out->z_b_lo))
            return PXH_E_ARG;
    } else {
        if (!fields_ok(old, true) || !lo_ok(g, {bed, old})) return PXH_E_ARG;
        const void *po[] = {old->h, old->hu, old->hv, old->z_b, old->G};
        for (const void *q : po)
            if (q == hyd->h || q == hyd->hu || q == hyd->hv || q == bed->z_b || q == bed->G) return PXH_E_ARG;
        if (old->z_b_lo && old->z_b_lo == bed->z_b_lo) return PXH_E_ARG;
    }
    // records as the fused step has them: first stage hyd = out records (or
    // neither), bed plain; second stage hyd plain, bed records or plain
    const bool ph = packed_view(g, hyd), pb = packed_view(g, bed);
    if (!old && (pb || ph != packed_view(g, out))) return PXH_E_ARG;
    if (old && (ph || packed_view(g, old))) return PXH_E_ARG;
    const int write_G = (flags & PXH_SWEEP_SKIP_G) && sd->mpm ? 0 : 1;
    return DISPATCH(g, rx_frame_sweep, g, hyd, bed, old ? nullptr : out, old, m0, sd, dt_dev, write_G, cfl,
                    old ? dt_partial : nullptr, static_cast<hipStream_t>(stream), ph ? 4 : 1, pb ? 4 : 1,
                    !old && ph ? 4 : 1);
}

int64_t pxh_multi_stage_blocks(const pxh_grid *g) {
    if (!grid_ok(g)) return PXH_E_ARG;
    if (!stage_fits(g)) return PXH_E_RANGE;
    const dim3 fg = stage_grid(g->ny, g->nx, strip_rows(g->flags));
    return int64_t(fg.x) * fg.y;
}

int64_t pxh_multi_frame_blocks(int32_t ny, int32_t nx) {
    if (ny < 2 || nx < 2) return PXH_E_ARG;
    return (frame_cells(ny, nx) + 255) / 256;
}

int64_t pxh_multi_seam_elems(int32_t ny, int32_t nx) {
    if (ny < 2 || nx < 2) return PXH_E_ARG;
    return seam_col_elems(ny, nx);
}

int64_t pxh_multi_dt_frame_blocks(const pxh_grid *g) {
    if (!grid_ok(g)) return PXH_E_ARG;
    if (!stage_fits(g)) return PXH_E_RANGE;
    return (band_dt_work(g->ny, g->nx, strip_rows(g->flags)) + 255) / 256;
}

int pxh_roe_exner_stage_sweep_multi(int32_t dtype, int32_t flags, const pxh_sweep_domain *doms, int32_t n,
                                    int32_t nblocks, int32_t second, int32_t lo, const pxh_sediments *sd,
                                    const void *dt_dev, int sweep_flags, void *dt_partial, void *xs, void *stream) {
    const int rows = PXH_STRIP_ROWS(flags);
    if ((dtype != PXH_F32 && dtype != PXH_F64) || !doms || n <= 0 || nblocks <= 0 || !sd || !dt_dev ||
        (rows != 0 && (rows < 8 || rows > 128)) || (!dt_partial != !xs) || (dt_partial && !second))
        return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    const int write_G = (sweep_flags & PXH_SWEEP_SKIP_G) && sd->mpm ? 0 : 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int sr = strip_rows(flags);
    if (dtype == PXH_F64)
        multi_stage_t<double, false>(doms, n, nblocks, sr, second, false, flags, static_cast<const double *>(dt_dev),
                                     sed_k(*sd), write_G, static_cast<double *>(dt_partial),
                                     static_cast<double *>(xs), st);
    else if (flags & PXH_EXACT_F32)
        multi_stage_t<float, false>(doms, n, nblocks, sr, second, lo, flags, static_cast<const float *>(dt_dev),
                                    sed_k(*sd), write_G, static_cast<float *>(dt_partial), static_cast<float *>(xs),
                                    st);
    else
        multi_stage_t<float, true>(doms, n, nblocks, sr, second, lo, flags, static_cast<const float *>(dt_dev),
                                   sed_k(*sd), write_G, static_cast<float *>(dt_partial), static_cast<float *>(xs),
                                   st);
    return launch_status();
}

int pxh_dt_frame_multi(int32_t dtype, int32_t flags, const pxh_sweep_domain *doms, int32_t n, int32_t nblocks,
                       int32_t nstage, double cfl, void *dt_partial, const void *xs, void *dt_out, void *stream) {
    const int rows = PXH_STRIP_ROWS(flags);
    if ((dtype != PXH_F32 && dtype != PXH_F64) || !doms || n <= 0 || nblocks <= 0 || nstage < 0 || !dt_partial ||
        !xs || !dt_out || (rows != 0 && (rows < 8 || rows > 128)) || int64_t(nstage) + nblocks >= (1LL << 31))
        return PXH_E_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int sr = strip_rows(flags);
    if (dtype == PXH_F64)
        return dt_frame_multi_t<double, false>(doms, n, nblocks, nstage, sr, cfl, static_cast<double *>(dt_partial),
                                               static_cast<const double *>(xs), static_cast<double *>(dt_out), st);
    if (flags & PXH_EXACT_F32)
        return dt_frame_multi_t<float, false>(doms, n, nblocks, nstage, sr, static_cast<float>(cfl),
                                              static_cast<float *>(dt_partial), static_cast<const float *>(xs),
                                              static_cast<float *>(dt_out), st);
    return dt_frame_multi_t<float, true>(doms, n, nblocks, nstage, sr, static_cast<float>(cfl),
                                         static_cast<float *>(dt_partial), static_cast<const float *>(xs),
                                         static_cast<float *>(dt_out), st);
}

int pxh_roe_exner_frame_sweep_multi(int32_t dtype, int32_t flags, const pxh_sweep_domain *doms, int32_t n,
                                    int32_t nblocks, int32_t second, const pxh_sediments *sd, const void *dt_dev,
                                    int sweep_flags, void *stream) {
    if ((dtype != PXH_F32 && dtype != PXH_F64) || !doms || n <= 0 || nblocks <= 0 || !sd || !dt_dev)
        return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    const int write_G = (sweep_flags & PXH_SWEEP_SKIP_G) && sd->mpm ? 0 : 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dtype == PXH_F64)
        multi_frame_t<double, false>(doms, n, nblocks, second, static_cast<const double *>(dt_dev), sed_k(*sd), write_G,
                                     st);
    else if (flags & PXH_EXACT_F32)
        multi_frame_t<float, false>(doms, n, nblocks, second, static_cast<const float *>(dt_dev), sed_k(*sd), write_G,
                                    st);
    else
        multi_frame_t<float, true>(doms, n, nblocks, second, static_cast<const float *>(dt_dev), sed_k(*sd), write_G,
                                   st);
    return launch_status();
}

int64_t pxh_dt_multi_blocks(int32_t ny, int32_t nx) {
    if (ny < 2 || nx < 2) return PXH_E_ARG;
    const dim3 sg = strip_grid<SO_WIDE, SR_BED>(ny, nx);
    return int64_t(sg.x) * sg.y;
}

int pxh_dt_multi(int32_t dtype, int32_t flags, const pxh_dt_domain *doms, int32_t n, int32_t nblocks, double cfl,
                 void *dt_out, void *partial, void *stream) {
    if ((dtype != PXH_F32 && dtype != PXH_F64) || n <= 0 || nblocks <= 0 || !doms || !dt_out || !partial)
        return PXH_E_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const dim3 sb(SW * SWPB);
    if (dtype == PXH_F64) {
        hipLaunchKernelGGL((k_multi_dt_strip<double, false>), dim3(nblocks), sb, 0, st, doms, int(n),
                           static_cast<double *>(partial));
        if (int e = launch_status()) return e;
        hipLaunchKernelGGL(k_dt_final<double>, dim3(1), dim3(FIN_THREADS), 0, st, static_cast<const double *>(partial),
                           int(nblocks), cfl, 1, static_cast<double *>(dt_out));
    } else {
        if (flags & PXH_EXACT_F32)
            hipLaunchKernelGGL((k_multi_dt_strip<float, false>), dim3(nblocks), sb, 0, st, doms, int(n),
                               static_cast<float *>(partial));
        else
            hipLaunchKernelGGL((k_multi_dt_strip<float, true>), dim3(nblocks), sb, 0, st, doms, int(n),
                               static_cast<float *>(partial));
        if (int e = launch_status()) return e;
        hipLaunchKernelGGL(k_dt_final<float>, dim3(1), dim3(FIN_THREADS), 0, st, static_cast<const float *>(partial),
                           int(nblocks), static_cast<float>(cfl), 1, static_cast<float *>(dt_out));
    }
    return launch_status();
}

int pxh_copy_batch(int32_t dtype, const pxh_copy_job *jobs, int32_t n, int64_t total, void *stream) {
    if ((dtype != PXH_F32 && dtype != PXH_F64) || n < 0 || total < 0 || (n > 0 && !jobs) || (total > 0 && n == 0))
        return PXH_E_ARG;
    if (total == 0) return PXH_OK;
    const unsigned nb = unsigned(total / 256 + 1 < 4096 ? total / 256 + 1 : 4096);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dtype == PXH_F64) hipLaunchKernelGGL(k_copy_batch<double>, dim3(nb), dim3(256), 0, st, jobs, int(n), long(total));
    else hipLaunchKernelGGL(k_copy_batch<float>, dim3(nb), dim3(256), 0, st, jobs, int(n), long(total));
    return launch_status();
}

int64_t pxh_dt_workspace_size(const pxh_grid *g) {
    if (!grid_ok(g)) return PXH_E_ARG;
    const int64_t esz = g->dtype == PXH_F64 ? 8 : 4;
    return dt_part_slots(g) * 8 + seam_col_elems(g->ny, g->nx) * esz;
}

int pxh_roe_exner_ssprk2(const pxh_grid *g, const pxh_fields *a, const pxh_fields *b, const pxh_fields *c,
                         const uint8_t *m0, const pxh_sediments *sd, const pxh_bc *bc, double cfl, void *dt_dev,
                         void *ws, int dt_ready, void *stream) {
    if (!grid_ok(g) || !fields_ok(a, true) || !fields_ok(b, true) || !fields_ok(c, true) || !m0 || !sd || !dt_dev ||
        !ws)
        return PXH_E_ARG;
    if (sd->n_sed < 0 || sd->n_sed > PXH_MAX_SED || (sd->mpm && sd->n_sed < 1)) return PXH_E_RANGE;
    const void *pa[] = {a->h, a->hu, a->hv, a->z_b, a->G}, *pb[] = {b->h, b->hu, b->hv, b->z_b, b->G},
               *pc[] = {c->h, c->hu, c->hv, c->z_b, c->G};
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            if (pa[i] == pb[j] || pa[i] == pc[j] || pb[i] == pc[j]) return PXH_E_ARG;
    if (!lo_ok(g, {a, b, c}) || (a->z_b_lo && (a->z_b_lo == b->z_b_lo || a->z_b_lo == c->z_b_lo || b->z_b_lo == c->z_b_lo)))
        return PXH_E_ARG;
    if (!bc_ok(bc)) return PXH_E_ARG;
    if (bc && bc->n_levels > 1) return PXH_E_UNSUPPORTED;
    return DISPATCH(g, rx_ssprk2, g, a, b, c, m0, sd, bc, cfl, dt_dev, ws, dt_ready,
                    static_cast<hipStream_t>(stream));
}

int pxh_roe_exner_ssprk2_noalloc(const pxh_grid *g, const pxh_fields *a, const pxh_fields *b, const pxh_fields *c,
                                 const uint8_t *m0, const pxh_sediments *sd, const pxh_bc *bc, double cfl,
                                 void *dt_dev, int dt_ready, void *stream) {
    if (!grid_ok(g)) return PXH_E_ARG;
    const int64_t need = pxh_dt_workspace_size(g);
    if (need < 0) return PXH_E_ARG;
    if (need > PXH_NOALLOC_SCRATCH_BYTES) return PXH_E_RANGE;
    void *ws = nullptr;
    if (int e = scratch_of(static_cast<hipStream_t>(stream), &ws)) return e;
    return pxh_roe_exner_ssprk2(g, a, b, c, m0, sd, bc, cfl, dt_dev, ws, dt_ready, stream);
}

int pxh_roe_exner_dt(const pxh_grid *g, const pxh_fields *s, const uint8_t *m0, double cfl, void *dt_out,
                     void *ws, void *stream) {
    if (!grid_ok(g) || !fields_ok(s, true) || !m0 || !dt_out || !ws) return PXH_E_ARG;
    return DISPATCH(g, rx_dt, g, s, m0, cfl, dt_out, ws, static_cast<hipStream_t>(stream));
}

int pxh_ssprk2_combine(const pxh_grid *g, const pxh_fields *a, const pxh_fields *b, const pxh_fields *o,
                       void *stream) {
    if (!grid_ok(g) || !fields_ok(a, true) || !fields_ok(b, true) || !fields_ok(o, true)) return PXH_E_ARG;
    if (!lo_ok(g, {a, b, o})) return PXH_E_ARG;
    const long n = long(g->ny) * g->nx;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (g->dtype == PXH_F64)
        hipLaunchKernelGGL(k_combine<double>, dim3((n + 255) / 256), dim3(256), 0, st, cast_fields<double>(a),
                           cast_fields<double>(b), cast_fields<double>(o), n);
    else
        hipLaunchKernelGGL(k_combine<float>, dim3((n + 255) / 256), dim3(256), 0, st, cast_fields<float>(a),
                           cast_fields<float>(b), cast_fields<float>(o), n);
    return launch_status();
}

int pxh_roe_step(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const pxh_bc *bc,
                 const void *dt_dev, void *stream) {
    if (!grid_ok(g) || !fields_ok(in, false) || !fields_ok(out, false) || !dt_dev) return PXH_E_ARG;
    if (in->h == out->h || in->hu == out->hu || in->hv == out->hv) return PXH_E_ARG;
    if (!bc_ok(bc)) return PXH_E_ARG;
    if (bc && bc->n_levels > 1) return PXH_E_UNSUPPORTED;
    return DISPATCH(g, roe_step, g, in, out, bc, dt_dev, static_cast<hipStream_t>(stream));
}

int pxh_roe_dt(const pxh_grid *g, const pxh_fields *s, double cfl, void *dt_out, void *ws, void *stream) {
    if (!grid_ok(g) || !fields_ok(s, false) || !dt_out || !ws) return PXH_E_ARG;
    return DISPATCH(g, roe_dt, g, s, cfl, dt_out, ws, static_cast<hipStream_t>(stream));
}

int pxh_bc_apply(const pxh_grid *g, const pxh_fields *s, const pxh_bc *bc, void *stream) {
    if (!grid_ok(g) || !s || !bc) return PXH_E_ARG;
    if (!bc_ok(bc)) return PXH_E_ARG;
    return DISPATCH(g, bc_apply_t, g, s, bc, static_cast<hipStream_t>(stream), packed_view(g, s) ? 4 : 1);
}

int64_t pxh_halo_bytes(const pxh_grid *g, const pxh_fields *s, int hw, uint32_t groups, int sides) {
    if (!grid_ok(g) || !s) return PXH_E_ARG;
    HaloPlan plan;
    if (g->dtype == PXH_F64) {
        HaloFields<double> hf;
        if (!halo_plan<double>(g, s, hw, groups, sides, plan, hf)) return PXH_E_ARG;
        return plan.off[plan.nf] * 8;
    }
    HaloFields<float> hf;
    if (!halo_plan<float>(g, s, hw, groups, sides, plan, hf)) return PXH_E_ARG;
    return plan.off[plan.nf] * 4;
}

int pxh_halo_pack(const pxh_grid *g, const pxh_fields *s, int hw, uint32_t groups, int sides, void *buf,
                  void *stream) {
    if (!grid_ok(g) || !s || !buf) return PXH_E_ARG;
    return DISPATCH(g, halo_t, false, g, s, hw, groups, sides, buf, static_cast<hipStream_t>(stream));
}

int pxh_halo_unpack(const pxh_grid *g, const pxh_fields *s, int hw, uint32_t groups, int sides, const void *buf,
                    void *stream) {
    if (!grid_ok(g) || !s || !buf) return PXH_E_ARG;
    return DISPATCH(g, halo_t, true, g, s, hw, groups, sides, const_cast<void *>(buf),
                    static_cast<hipStream_t>(stream));
}

int64_t pxh_ensemble_plan(int32_t dtype, int32_t flags, const pxh_member *members, int32_t n, void *plan,
                          int64_t plan_bytes) {
    return ens_plan_build(dtype, flags, members, n, plan, plan_bytes);
}

int64_t pxh_ensemble_workspace_size(const void *plan) {
    const EnsPlanHeader *h = ens_header(plan);
    if (!h) return PXH_E_ARG;
    return ens_xs_offset(*h) + h->xs_elems * (h->dtype == PXH_F64 ? 8 : 4);
}

int pxh_ensemble_dt(const void *plan, const void *plan_dev, const pxh_fields *a, const uint8_t *m0, double cfl,
                    void *dt_dev, void *workspace, void *stream) {
    const EnsPlanHeader *h = ens_header(plan);
    if (!h || !plan_dev || !fields_ok(a, true) || !m0 || !dt_dev || !workspace) return PXH_E_ARG;
    const char *pd = static_cast<const char *>(plan_dev);
    hipStream_t st = static_cast<hipStream_t>(stream);
    return h->dtype == PXH_F64 ? ens_dt<double>(*h, pd, a, m0, cfl, dt_dev, workspace, st)
                               : ens_dt<float>(*h, pd, a, m0, cfl, dt_dev, workspace, st);
}

int pxh_ensemble_ssprk2(const void *plan, const void *plan_dev, const pxh_fields *a, const pxh_fields *b,
                        const pxh_fields *c, const uint8_t *m0, const pxh_bc *bc, double cfl, void *dt_dev,
                        void *workspace, int dt_ready, void *stream) {
    const EnsPlanHeader *h = ens_header(plan);
    if (!h || !plan_dev || !fields_ok(a, true) || !fields_ok(b, true) || !fields_ok(c, true) || !m0 || !dt_dev ||
        !workspace)
        return PXH_E_ARG;
    if (!bc_ok(bc)) return PXH_E_ARG;
    if (bc && bc->n_entries > 0 && (!(bc->flags & PXH_BC_FRAME) || bc->n_levels > 1)) return PXH_E_UNSUPPORTED;
    const pxh_grid g{1, 1, h->dtype, h->flags, 1.0};
    if (!lo_ok(&g, {a, b, c})) return PXH_E_ARG;
    const void *pa[] = {a->h, a->hu, a->hv, a->z_b, a->G}, *pb[] = {b->h, b->hu, b->hv, b->z_b, b->G},
               *pc[] = {c->h, c->hu, c->hv, c->z_b, c->G};
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            if (pa[i] == pb[j] || pa[i] == pc[j] || pb[i] == pc[j]) return PXH_E_ARG;
    const char *pd = static_cast<const char *>(plan_dev);
    hipStream_t st = static_cast<hipStream_t>(stream);
    return h->dtype == PXH_F64 ? ens_ssprk2<double>(*h, pd, a, b, c, m0, bc, cfl, dt_dev, workspace, dt_ready, st)
                               : ens_ssprk2<float>(*h, pd, a, b, c, m0, bc, cfl, dt_dev, workspace, dt_ready, st);
}

}  // extern "C"
