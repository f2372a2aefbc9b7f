// Batched Roe-Exner SSPRK2: many independent domains (an ensemble of
// scenarios of uneven size, each with its own sediment table and dt) advanced
// by one launch per sweep for all members.
//
// Every member's fields live in shared arenas (one per field), member m at
// element offset `off`.  A member table on the device gives each member's
// shape and the first block of its strips in each launch; a block finds its
// member by a binary search over those starts (wave-uniform, scalar loads) and
// then runs the same strip body as the single-domain kernels.  Launch for
// launch the batch is the fused SSPRK2 step of stage.hpp:
//   stage sweep (all members) -> ring fill -> frame sweep -> ring fill ->
//   stage sweep + combine + dt minima -> ring fill -> frame sweep + frame dt
//   -> ring fill -> one block per member reduces its minima into its dt.
#pragma once
#include "stage.hpp"

namespace pxh {

struct EnsMember {
    long off;      // first element of the member's arrays in the arenas
    long nband;    // cells of its two-cell frame
    long xs0;      // first element of its seam-column buffer (seam_rec)
    int ny, nx;
    int sgx;       // stage-grid columns
    int sblk0;     // first stage-sweep block
    int fblk0;     // first frame block (256 cells each)
    int dgx;       // standalone-dt grid columns
    int dblk0;     // first standalone-dt block
    int part0;     // first of its partial minima
    int nstage;    // stage-sweep blocks
    int nframe;    // frame blocks
    int ndt;       // standalone-dt blocks
    int tblk0;     // first frame + seam dt block
    int ntdt;      // frame + seam dt blocks
    int pad;
    double dx;
};

// The member whose block range [start, next start) holds blk; tab has n + 1
// entries, the last a sentinel with the launch's block count.
template <int EnsMember::*START>
__device__ __forceinline__ int find_member(const EnsMember *__restrict__ tab, int n, int blk) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tab[mid].*START <= blk) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// A member's fields at element offset off; rec: f's h, hu, hv, z_b are the
// four-element records of PXH_PACKED_STAGE over the arena, so the member's
// records start at f.h + 4 off, with hu, hv, z_b one, two, three elements on.
template <typename T>
__device__ __forceinline__ F<T> shifted_rec(const F<T> &f, long off, bool rec) {
    F<T> r = shifted(f, off);
    if (rec && f.h) {
        r.h = f.h + 4 * off;
        r.hu = r.h + 1;
        r.hv = r.h + 2;
        r.z_b = r.h + 3;
    }
    return r;
}

// ST: the static-field specialisation of stage_strip_run, a promise that
// holds for every member (each member's uniform z is its own first cell).
// PK: the first stage's output (out in MODE 0, in in MODE 1) as records.
template <typename T, bool FAST, int MODE, bool DT, bool LO, bool MPM, int ST, bool PK = false>
__global__ __launch_bounds__(SW *FWPB) void k_ens_stage(F<T> in, F<T> out, F<T> cb, F<T> old,
                                                        const uint8_t *__restrict__ m0, const T *__restrict__ dts,
                                                        const EnsMember *__restrict__ tab, int n, int sr,
                                                        const SedK *__restrict__ seds, int write_G,
                                                        T *__restrict__ partial, T *__restrict__ xs) {
    __shared__ T red[FWPB];
    int lb, unused;
    xcd_block(lb, unused);  // one-dimensional grid: lb is the XCD-ordered block
    const int i = find_member<&EnsMember::sblk0>(tab, n, lb);
    const EnsMember m = tab[i];
    const int local = lb - m.sblk0;
    const int by = local / m.sgx, bx = local - by * m.sgx;
    const StagePos p = stage_pos_at(bx, by, m.ny, m.nx, sr);
    T vmin = stage_strip_run<T, FAST, MODE, DT, LO, MPM, ST, PK>(
        p, shifted_rec(in, m.off, PK && MODE == 1), shifted_rec(out, m.off, PK && MODE == 0), shifted(cb, m.off),
        shifted(old, m.off), m0 + m.off,
                                                             dts + i, m.ny, m.nx, static_cast<T>(m.dx), seds[i],
                                                             write_G, DT ? xs + m.xs0 : nullptr);
    if (MODE == 1 && DT) {
        vmin = wave_block_min_n<FWPB>(vmin, red);
        if (threadIdx.x == 0) partial[m.part0 + local] = vmin;
    }
}

template <typename T, bool FAST, int MODE>
__global__ __launch_bounds__(256) void k_ens_band_bed(F<T> hyd, F<T> bed, F<T> out, F<T> old,
                                                      const uint8_t *__restrict__ m0, const T *__restrict__ dts,
                                                      const EnsMember *__restrict__ tab, int n,
                                                      const SedK *__restrict__ seds, int write_G, int hs, int zs,
                                                      int os) {
    const int i = find_member<&EnsMember::fblk0>(tab, n, blockIdx.x);
    const EnsMember m = tab[i];
    const long t = long(blockIdx.x - m.fblk0) * 256 + threadIdx.x;
    if (t < m.nband)
        band_bed_cell<T, FAST, MODE>(t, shifted_rec(hyd, m.off, hs == 4), shifted_rec(bed, m.off, zs == 4),
                                     shifted_rec(out, m.off, os == 4), shifted(old, m.off), m0 + m.off, dts + i,
                                     m.ny, m.nx, static_cast<T>(m.dx), seds[i], write_G, hs, zs, os);
}

// The dt minima of each member's frame edges and strip-seam edges.
template <typename T, bool FAST>
__global__ __launch_bounds__(256) void k_ens_band_dt(F<T> s, const uint8_t *__restrict__ m0,
                                                     const EnsMember *__restrict__ tab, int n, int sr,
                                                     const T *__restrict__ xs, T *__restrict__ partial) {
    const int i = find_member<&EnsMember::tblk0>(tab, n, blockIdx.x);
    const EnsMember m = tab[i];
    const int local = blockIdx.x - m.tblk0;
    const long t = long(local) * 256 + threadIdx.x;
    block256_min_to(band_dt_cell<T, FAST>(t, m.nband, band_dt_work(m.ny, m.nx, sr), shifted(s, m.off), m0 + m.off,
                                          xs + m.xs0, m.ny, m.nx, static_cast<T>(m.dx), sr),
                    partial + m.part0 + m.nstage + local);
}

template <typename T, bool FAST>
__global__ __launch_bounds__(SW *SWPB) void k_ens_dt_strip(F<T> s, const uint8_t *__restrict__ m0,
                                                           const EnsMember *__restrict__ tab, int n,
                                                           T *__restrict__ partial) {
    __shared__ T red[SWPB];
    int lb, unused;
    xcd_block(lb, unused);
    const int i = find_member<&EnsMember::dblk0>(tab, n, lb);
    const EnsMember m = tab[i];
    const int local = lb - m.dblk0;
    const int by = local / m.dgx, bx = local - by * m.dgx;
    const StripPos p = strip_pos_at<SO_WIDE, SR_BED>(bx, by, m.ny, m.nx);
    T v = dt_strip_run<T, FAST>(p, shifted(s, m.off), m0 + m.off, m.ny, m.nx, static_cast<T>(m.dx));
    v = wave_block_min(v, red);
    if (threadIdx.x == 0) partial[m.part0 + local] = v;
}

// Field pointers of a pxh_fields on the device.
template <typename T>
__device__ __forceinline__ F<T> dev_fields(const pxh_fields &f) {
    F<T> r;
    r.h = static_cast<T *>(f.h);
    r.hu = static_cast<T *>(f.hu);
    r.hv = static_cast<T *>(f.hv);
    r.z = static_cast<T *>(f.z);
    r.z_b = static_cast<T *>(f.z_b);
    r.n = static_cast<T *>(f.n);
    r.n_b = static_cast<T *>(f.n_b);
    r.G = static_cast<T *>(f.G);
    r.z_b_lo = sizeof(T) == 4 ? static_cast<T *>(f.z_b_lo) : nullptr;
    return r;
}

// The domain of a multi-domain launch whose block range holds blk.
template <int pxh_sweep_domain::*START>
__device__ __forceinline__ int find_domain(const pxh_sweep_domain *__restrict__ doms, int n, int blk) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (doms[mid].*START <= blk) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// The fused stage sweep over several domains (pxh_roe_exner_stage_sweep_multi);
// ST as for k_ens_stage, holding for every domain.  DT (second stage): each
// block's CFL minimum over its domain's interior edges into partial[block],
// each domain's seam records at xs + xs0.
template <typename T, bool FAST, int MODE, bool LO, bool MPM, int ST, bool DT>
__global__ __launch_bounds__(SW *FWPB) void k_multi_stage(const pxh_sweep_domain *__restrict__ doms, int n, int sr,
                                                          const T *__restrict__ dtp, SedK sd, int write_G,
                                                          T *__restrict__ partial, T *__restrict__ xs) {
    __shared__ T red[FWPB];
    int lb, unused;
    xcd_block(lb, unused);  // one-dimensional grid
    const pxh_sweep_domain &d = doms[find_domain<&pxh_sweep_domain::sblk0>(doms, n, lb)];
    const int ny = d.ny, nx = d.nx;
    const int sgx = stage_gx(nx);
    const int local = lb - d.sblk0;
    const int by = local / sgx, bx = local - by * sgx;
    const StagePos p = stage_pos_at(bx, by, ny, nx, sr);
    const F<T> none{};
    if constexpr (MODE == 0)
        stage_strip_run<T, FAST, 0, false, LO, MPM, ST>(p, dev_fields<T>(d.in), dev_fields<T>(d.out), none, none,
                                                        d.mask0, dtp, ny, nx, static_cast<T>(d.dx), sd, write_G,
                                                        nullptr);
    else {
        T vmin = stage_strip_run<T, FAST, 1, DT, LO, MPM, ST>(p, dev_fields<T>(d.in), none, dev_fields<T>(d.out),
                                                              dev_fields<T>(d.old), d.mask0, dtp, ny, nx,
                                                              static_cast<T>(d.dx), sd, write_G,
                                                              DT ? xs + d.xs0 : nullptr);
        if constexpr (DT) {
            vmin = wave_block_min_n<FWPB>(vmin, red);
            if (threadIdx.x == 0) partial[lb] = vmin;
        }
    }
}

// The frame and strip-seam edges of every domain's combined state (old) for
// pxh_dt_frame_multi, one partial minimum per block, folded with a share of
// the nprev per-block minima `prev` of the second-stage sweep (grid-strided),
// so the final reduction reads this kernel's partials only; frame edges
// touching a shared side's ghost rings are skipped (band_dt_cell).
template <typename T, bool FAST>
__global__ __launch_bounds__(256) void k_multi_band_dt(const pxh_sweep_domain *__restrict__ doms, int n, int sr,
                                                       const T *__restrict__ xs, const T *__restrict__ prev,
                                                       int nprev, T *__restrict__ partial) {
    const pxh_sweep_domain &d = doms[find_domain<&pxh_sweep_domain::tblk0>(doms, n, blockIdx.x)];
    const long t = long(blockIdx.x - d.tblk0) * 256 + threadIdx.x;
    T v = band_dt_cell<T, FAST>(t, frame_cells(d.ny, d.nx), band_dt_work(d.ny, d.nx, sr), dev_fields<T>(d.old),
                                d.mask0, xs + d.xs0, d.ny, d.nx, static_cast<T>(d.dx), sr, d.shared);
    for (long k = long(blockIdx.x) * 256 + threadIdx.x; k < nprev; k += long(gridDim.x) * 256) v = pmin(v, prev[k]);
    block256_min_to(v, partial + blockIdx.x);
}

// The frame sweep over several domains (pxh_roe_exner_frame_sweep_multi):
// hyd = out, bed = in.
template <typename T, bool FAST, int MODE>
__global__ __launch_bounds__(256) void k_multi_band_bed(const pxh_sweep_domain *__restrict__ doms, int n,
                                                        const T *__restrict__ dtp, SedK sd, int write_G) {
    const pxh_sweep_domain &d = doms[find_domain<&pxh_sweep_domain::fblk0>(doms, n, blockIdx.x)];
    const long t = long(blockIdx.x - d.fblk0) * 256 + threadIdx.x;
    if (t < frame_cells(d.ny, d.nx)) {
        const F<T> none{};
        const F<T> out = dev_fields<T>(d.out);
        band_bed_cell<T, FAST, MODE>(t, out, dev_fields<T>(d.in), MODE == 0 ? out : none,
                                     MODE == 1 ? dev_fields<T>(d.old) : none, d.mask0, dtp, d.ny, d.nx,
                                     static_cast<T>(d.dx), sd, write_G);
    }
}

// The standalone dt sweep over several domains of their own allocations
// (pxh_dt_multi: the tiles of one device), one partial per block; domain k's
// blocks start at doms[k].blk0.
template <typename T, bool FAST>
__global__ __launch_bounds__(SW *SWPB) void k_multi_dt_strip(const pxh_dt_domain *__restrict__ doms, int n,
                                                             T *__restrict__ partial) {
    __shared__ T red[SWPB];
    int lb, unused;
    xcd_block(lb, unused);  // one-dimensional grid
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (doms[mid].blk0 <= lb) lo = mid;
        else hi = mid;
    }
    const pxh_dt_domain d = doms[__builtin_amdgcn_readfirstlane(lo)];
    const int gx = (d.nx + SO_WIDE - 1) / SO_WIDE;
    const int local = lb - d.blk0;
    const int by = local / gx, bx = local - by * gx;
    const StripPos p = strip_pos_at<SO_WIDE, SR_BED>(bx, by, d.ny, d.nx);
    F<T> s{};
    s.h = static_cast<T *>(const_cast<void *>(d.h));
    s.hu = static_cast<T *>(const_cast<void *>(d.hu));
    s.hv = static_cast<T *>(const_cast<void *>(d.hv));
    s.G = static_cast<T *>(const_cast<void *>(d.G));
    T v = dt_strip_run<T, FAST>(p, s, d.mask0, d.ny, d.nx, static_cast<T>(d.dx));
    v = wave_block_min(v, red);
    if (threadIdx.x == 0) partial[lb] = v;
}

// One block per member: dts[i] = max(cfl * min of its minima, 1e-6)
// (compute_dt_roeexner, roe_exner_solver.py:118-122).  step: the minima of a
// fused SSPRK2 step (stage + frame/seam blocks); else of the standalone dt sweep.
template <typename T>
__global__ __launch_bounds__(256) void k_ens_dt_final(const T *__restrict__ partial,
                                                      const EnsMember *__restrict__ tab, T cfl, int step,
                                                      T *__restrict__ dts) {
    __shared__ T red[4];
    const EnsMember m = tab[blockIdx.x];
    const int cnt = step ? m.nstage + m.ntdt : m.ndt;
    T v = inf_v<T>();
    for (int k = threadIdx.x; k < cnt; k += 256) v = pmin(v, partial[m.part0 + k]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) v = pmin(v, red[w]);
        dts[blockIdx.x] = pmax(cfl * v, C<T>(1e-6));
    }
}

}  // namespace pxh
