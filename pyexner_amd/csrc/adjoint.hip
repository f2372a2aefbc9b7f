// Adjoint entry points of the Roe-Exner step (double only): one stage, the
// CFL dt, one SSPRK2 step and the initialisation, each the transpose of the
// forward entry of the same name.  The kernels are in adjoint.hpp.
#include <hip/hip_runtime.h>

#include "adjoint.hpp"

namespace pxh {
namespace {

inline dim3 adj_grid(int ny, int nx) { return dim3((nx + ADJ_BX - 1) / ADJ_BX, (ny + ADJ_BY - 1) / ADJ_BY); }
inline int adj_blocks(int ny, int nx) {
    const dim3 g = adj_grid(ny, nx);
    return int(g.x * g.y);
}
inline long align_d(long n) { return (n + 31) & ~31L; }  // doubles, to 256 bytes

// The workspace: Lc and Lb (five fields each), h1, partial sums (d/d dt in
// four slots of dt_seg, the parameters in two of par_seg: an SSPRK2 step
// keeps each kernel's partials and sums them once), the dt argmin partials
// and two scalars.
struct AdjWs {
    F<double> lc, lb;
    double *h1, *part_dt, *part_par, *arg_v, *lam_dt;
    long *arg_i;
    long dt_seg, par_seg;
};

inline long ws_doubles(int ny, int nx) {
    const long n = long(ny) * nx, nc = align_d(n), nb = adj_blocks(ny, nx), na = (n + 255) / 256;
    return 11 * nc + 4 * align_d(nb) + 2 * align_d(ADJ_NPAR * nb) + 2 * align_d(na) + 32;
}

inline AdjWs ws_carve(int ny, int nx, void *ws) {
    const long n = long(ny) * nx, nc = align_d(n), nb = adj_blocks(ny, nx), na = (n + 255) / 256;
    double *p = static_cast<double *>(ws);
    auto take = [&](long k) {
        double *r = p;
        p += k;
        return r;
    };
    AdjWs w{};
    w.lc.h = take(nc); w.lc.hu = take(nc); w.lc.hv = take(nc); w.lc.z_b = take(nc); w.lc.G = take(nc);
    w.lb.h = take(nc); w.lb.hu = take(nc); w.lb.hv = take(nc); w.lb.z_b = take(nc); w.lb.G = take(nc);
    w.h1 = take(nc);
    w.dt_seg = align_d(nb);
    w.par_seg = align_d(ADJ_NPAR * nb);
    w.part_dt = take(4 * w.dt_seg);
    w.part_par = take(2 * w.par_seg);
    w.arg_v = take(align_d(na));
    w.arg_i = reinterpret_cast<long *>(take(align_d(na)));
    w.lam_dt = take(32);
    return w;
}

inline bool adj_grid_ok(const pxh_grid *g) { return grid_ok(g) && g->dtype == PXH_F64; }

// The outflux depth (a Newton solve from the source cell) has no adjoint here,
// nor has a program of several dependency levels (the transposed fill runs as
// one level).
inline bool bc_adj_ok(const pxh_bc *bc) {
    if (!bc) return true;
    if (bc->n_levels > 1) return false;
    for (int k = 0; k < bc->n_values; ++k)
        if (bc->values[k].kind == PXH_BV_OUTFLUX) return false;
    return true;
}

inline bool lam_ok(const pxh_fields *l, bool zb, bool G) {
    return l && l->h && l->hu && l->hv && (!zb || l->z_b) && (!G || l->G);
}

int bc_t(const F<double> &lam, const pxh_bc *bc, uint32_t groups, hipStream_t st) {
    if (!bc || bc->n_entries <= 0) return PXH_OK;
    const uint32_t gm = bc->group_mask & groups;
    if (!gm) return PXH_OK;
    hipLaunchKernelGGL(k_adj_bc_move, dim3((bc->n_entries + 255) / 256), dim3(256), 0, st, lam, bc->entries,
                       bc->n_entries, *bc, gm, (bc->flags & PXH_BC_SRC_SORTED) ? 1 : 0);
    return launch_status();
}

constexpr uint32_t HY = (1u << PXH_G_H) | (1u << PXH_G_HUV);
constexpr uint32_t STEP = HY | (1u << PXH_G_ZB) | (1u << PXH_G_G);

// The stage adjoint.  in: the stage input (h, hu, hv, z, z_b, n_b); out: its
// output's filled hydro (h, hu, hv); lo: the output's adjoint (h, hu, hv, z_b,
// G; consumed); li: accumulates the input's adjoint (h, hu, hv, z_b, and n_b
// when set); lam_dt and lam_par accumulate.  slot < 0: the partial sums are
// reduced into lam_dt and lam_par here; slot 0 or 1: they are left in
// part_dt slots 2 slot and 2 slot + 1 and part_par slot `slot` for the
// caller to reduce (lam_dt and lam_par untouched).
int stage_t(const pxh_grid *g, const F<double> &in, const F<double> &out, const uint8_t *m0,
            const pxh_sediments *sd, const pxh_bc *bc, const double *dtp, const F<double> &lo, const F<double> &li,
            double *lam_dt, double *lam_par, const AdjWs &w, int slot, hipStream_t st) {
    const int ny = g->ny, nx = g->nx, nb = adj_blocks(ny, nx);
    const double dx = g->dx;
    const dim3 ag = adj_grid(ny, nx), ab(ADJ_BX, ADJ_BY);
    const int par = lam_par != nullptr;
    const int s = slot < 0 ? 0 : slot;
    double *pdt_ex = w.part_dt + 2 * s * w.dt_seg, *pdt_hy = pdt_ex + w.dt_seg, *ppar = w.part_par + s * w.par_seg;
    // final fill of (h, hu, hv, z_b, G)
    if (int e = bc_t(lo, bc, STEP, st)) return e;
    // Exner + compute_G: lam of the filled hydro accumulates in lo's h, hu, hv
    hipLaunchKernelGGL(k_adj_exner, ag, ab, 0, st, out, in.z_b, m0, dtp, ny, nx, dx, *sd, lo, lo, li.z_b, pdt_ex,
                       ppar, par);
    if (int e = launch_status()) return e;
    if (slot < 0) {
        hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, st, pdt_ex, long(nb), 1, 0, lam_dt, 1, 1, 0L);
        if (par) hipLaunchKernelGGL(k_adj_sum, dim3(4), dim3(256), 0, st, ppar, long(nb), 4, 0, lam_par, 1, 1, 0L);
        if (int e = launch_status()) return e;
    }
    // fill of (h, hu, hv) after the hydro update
    if (int e = bc_t(lo, bc, HY, st)) return e;
    // momentum corrections, on the raw hydro depth
    if (int e = hipMemcpyAsync(w.h1, out.h, sizeof(double) * size_t(ny) * nx, hipMemcpyDeviceToDevice, st))
        return e;
    if (bc && bc->n_entries > 0) {
        hipLaunchKernelGGL(k_adj_h1_targets, dim3((bc->n_entries + 255) / 256), dim3(256), 0, st, in, m0, dtp, ny,
                           nx, dx, bc->entries, bc->n_entries, w.h1);
        if (int e = launch_status()) return e;
    }
    hipLaunchKernelGGL(k_adj_momcorr, ag, ab, 0, st, w.h1, in.z, in.z_b, m0, ny, nx, lo);
    if (int e = launch_status()) return e;
    // hydro update
    hipLaunchKernelGGL(k_adj_hydro, ag, ab, 0, st, in, m0, dtp, ny, nx, dx, lo, li, pdt_hy);
    if (int e = launch_status()) return e;
    if (slot < 0) hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, st, pdt_hy, long(nb), 1, 0, lam_dt, 1, 1, 0L);
    return launch_status();
}

int dt_t(const pxh_grid *g, const F<double> &s, const uint8_t *m0, double cfl, const double *lam_dt,
         const F<double> &ls, const AdjWs &w, hipStream_t st) {
    const int ny = g->ny, nx = g->nx;
    const int na = int((long(ny) * nx + 255) / 256);
    hipLaunchKernelGGL(k_adj_dt_argmin, dim3(na), dim3(256), 0, st, s, m0, ny, nx, g->dx, w.arg_v, w.arg_i);
    if (int e = launch_status()) return e;
    hipLaunchKernelGGL(k_adj_dt_apply, dim3(1), dim3(256), 0, st, s, nx, g->dx, cfl, w.arg_v, w.arg_i, na, lam_dt,
                       ls);
    return launch_status();
}

}  // namespace
}  // namespace pxh

using namespace pxh;

extern "C" {

int64_t pxh_adj_workspace_size(const pxh_grid *g, const pxh_bc *largest_bc) {
    if (!adj_grid_ok(g) || !bc_ok(largest_bc)) return PXH_E_ARG;
    return int64_t(sizeof(double)) * ws_doubles(g->ny, g->nx);
}

int pxh_adj_roe_exner_stage(const pxh_grid *g, const pxh_fields *in, const pxh_fields *out, const uint8_t *m0,
                            const pxh_sediments *sd, const pxh_bc *bc, const void *dt_dev, const pxh_fields *lam_out,
                            const pxh_fields *lam_in, void *lam_dt, void *lam_par, void *ws, void *stream) {
    if (!adj_grid_ok(g) || !fields_ok(in, true) || !lam_ok(out, false, false) || !m0 || !sd || !dt_dev ||
        !lam_ok(lam_out, true, true) || !lam_ok(lam_in, true, false) || !lam_dt || !ws || !bc_ok(bc))
        return PXH_E_ARG;
    if (!sd->mpm || !bc_adj_ok(bc)) return PXH_E_UNSUPPORTED;
    const AdjWs w = ws_carve(g->ny, g->nx, ws);
    return stage_t(g, cast_fields<double>(in), cast_fields<double>(out), m0, sd, bc,
                   static_cast<const double *>(dt_dev), cast_fields<double>(lam_out), cast_fields<double>(lam_in),
                   static_cast<double *>(lam_dt), static_cast<double *>(lam_par), w, -1, hipStream_t(stream));
}

int pxh_adj_roe_exner_dt(const pxh_grid *g, const pxh_fields *s, const uint8_t *m0, double cfl, const void *lam_dt,
                         const pxh_fields *lam_s, void *ws, void *stream) {
    if (!adj_grid_ok(g) || !fields_ok(s, true) || !m0 || !lam_dt || !lam_ok(lam_s, false, true) || !ws)
        return PXH_E_ARG;
    const AdjWs w = ws_carve(g->ny, g->nx, ws);
    return dt_t(g, cast_fields<double>(s), m0, cfl, static_cast<const double *>(lam_dt), cast_fields<double>(lam_s),
                w, hipStream_t(stream));
}

int pxh_adj_roe_exner_ssprk2(const pxh_grid *g, const pxh_fields *a_, const pxh_fields *b_, const pxh_fields *c_,
                             const uint8_t *m0, const pxh_sediments *sd, const pxh_bc *bc, double cfl,
                             const void *dt_dev, const pxh_fields *lam, void *lam_par, void *ws, void *stream) {
    if (!adj_grid_ok(g) || !fields_ok(a_, true) || !lam_ok(b_, true, false) || !lam_ok(c_, false, false) || !m0 ||
        !sd || !dt_dev || !lam_ok(lam, true, true) || !ws || !bc_ok(bc))
        return PXH_E_ARG;
    if (!sd->mpm || !bc_adj_ok(bc)) return PXH_E_UNSUPPORTED;
    hipStream_t st = hipStream_t(stream);
    const AdjWs w = ws_carve(g->ny, g->nx, ws);
    const F<double> a = cast_fields<double>(a_), la = cast_fields<double>(lam);
    // b: stage-1 output and stage-2 input (its static fields are a's)
    F<double> b = a;
    b.h = static_cast<double *>(b_->h); b.hu = static_cast<double *>(b_->hu); b.hv = static_cast<double *>(b_->hv);
    b.z_b = static_cast<double *>(b_->z_b);
    const F<double> c = cast_fields<double>(c_);
    const double *dtp = static_cast<const double *>(dt_dev);
    const long n = long(g->ny) * g->nx;
    const int nb = adj_blocks(g->ny, g->nx);
    hipLaunchKernelGGL(k_adj_split, dim3((n + 255) / 256), dim3(256), 0, st, la, w.lc, w.lb, n);
    if (int e = launch_status()) return e;
    F<double> lb_in = w.lb;
    lb_in.n_b = la.n_b;
    double *lp = static_cast<double *>(lam_par);
    // the two stages leave their partial sums in slots 1 and 0, reduced once
    // here: d/d dt over all four kernels into lam_dt (overwritten), the
    // parameters of both Exner kernels added to lam_par
    if (int e = stage_t(g, b, c, m0, sd, bc, dtp, w.lc, lb_in, w.lam_dt, lp, w, 1, st)) return e;
    if (int e = stage_t(g, a, b, m0, sd, bc, dtp, w.lb, la, w.lam_dt, lp, w, 0, st)) return e;
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, st, w.part_dt, long(nb), 1, 0, w.lam_dt, 0, 4, w.dt_seg);
    if (lp) hipLaunchKernelGGL(k_adj_sum, dim3(4), dim3(256), 0, st, w.part_par, long(nb), 4, 0, lp, 1, 2, w.par_seg);
    if (int e = launch_status()) return e;
    return dt_t(g, a, m0, cfl, w.lam_dt, la, w, st);
}

int pxh_adj_roe_exner_init(const pxh_grid *g, const pxh_fields *s, const uint8_t *m0, const pxh_sediments *sd,
                           const pxh_bc *bc, const pxh_fields *lam, void *lam_par, void *ws, void *stream) {
    if (!adj_grid_ok(g) || !fields_ok(s, true) || !m0 || !sd || !lam_ok(lam, true, true) || !ws || !bc_ok(bc))
        return PXH_E_ARG;
    if (!sd->mpm || !bc_adj_ok(bc)) return PXH_E_UNSUPPORTED;
    if (lam_par && !lam->n_b) return PXH_E_ARG;
    hipStream_t st = hipStream_t(stream);
    const AdjWs w = ws_carve(g->ny, g->nx, ws);
    const F<double> la = cast_fields<double>(lam);
    if (int e = bc_t(la, bc, 0xffffffffu, st)) return e;
    const int ny = g->ny, nx = g->nx, nb = adj_blocks(ny, nx);
    const int par = lam_par != nullptr;
    hipLaunchKernelGGL(k_adj_init, adj_grid(ny, nx), dim3(ADJ_BX, ADJ_BY), 0, st, cast_fields<double>(s), m0, ny, nx,
                       *sd, la, w.part_par, par);
    if (int e = launch_status()) return e;
    if (par)
        hipLaunchKernelGGL(k_adj_sum, dim3(ADJ_NPAR), dim3(256), 0, st, w.part_par, long(nb), ADJ_NPAR, 0,
                           static_cast<double *>(lam_par), 1, 1, 0L);
    return launch_status();
}

}  // extern "C"
