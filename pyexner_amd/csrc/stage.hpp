// Fused Roe-Exner stage: roe_solve_2D + momentum_corrections + compute_G +
// exner_solve_2D (+ the SSPRK2 combine and the next dt's per-block minima) in
// one pass over the state, for boundary programs whose ring fill stays in the
// frame of the grid.
//
// A stage of the reference (roe_exner_solver.py:80-116) fills the ghost ring
// between the hydro update and the bed update.  When every target of the fill
// lies on the outer ring of cells and every source within two cells of the
// grid edge (Reflective, Transmissive and Periodic rings do: their sources sit
// in rows/columns 1 and n-2), only cells of the two-cell frame read filled values in the bed
// update.  The fused kernel then updates the bed of every cell off the frame
// right behind the hydro update, from the updated rows it holds in registers,
// and the frame follows after the fill in small band kernels (k_rx_band_*).
//
// Strip geometry: lanes 2..61 own 60 columns; the hydro update is complete on
// lanes 1..62 (both of their x edges are in the wavefront), which the bed
// edges of the owned columns need.  A strip of SR_STAGE owned rows updates the
// hydro of one row above and one below it.
#pragma once
#include "strip.hpp"

namespace pxh {

#ifndef PXH_SR_STAGE
#define PXH_SR_STAGE 32
#endif
#ifndef PXH_FWPB
#define PXH_FWPB 1
#endif
constexpr int FO = 60;          // owned columns per fused strip
constexpr int FHW = 2;          // halo lanes on each side
constexpr int SR_STAGE = PXH_SR_STAGE;  // owned rows per fused strip unless the grid's flags choose
// Fused strips (wavefronts) per block, side by side in x; with more than one
// the waves of a block walk the same rows in lockstep (a barrier per row), so
// that the 64-byte blocks two neighbouring strips share at their common
// column edge are written by both at nearly the same time.  A probe of the
// stage's access pattern alone (7 loads and 5 stores per cell, 60-column
// strips, 8192 x 8192) took 820 us with one wave per block, 704 us with 4
// waves in lockstep and 673 us with 16 (a flat stream of the same bytes:
// 561 us).  The stage kernels themselves do not gain: 4 waves per block gave
// 767 / 1216 us against 788 / 1193 us with one (the probe of the second
// stage's pattern: 1070 us in lockstep blocks of 4 against 1324 us), so the
// sweeps are held by their arithmetic's latency once the pattern improves;
// 2 waves per block gave the same step time.
constexpr int FWPB = PXH_FWPB;
constexpr int SR_MAX = 128;     // the largest strip height a grid may choose

struct StagePos {
    int lane, c, cc, r0, r1;
    int bx;     // strip column
    bool live;  // the strip has columns in the grid (the last block's extra waves do not)
    bool out;   // owned column
    bool hval;  // updated depth complete in this lane
};

// Owned rows per fused strip of a grid (PXH_STRIP_ROWS of its flags).
__host__ __device__ inline int strip_rows(int flags) {
    const int r = PXH_STRIP_ROWS(flags);
    return r ? r : SR_STAGE;
}

// This wave's strip of block (bx, by) of a domain's stage grid of sr-row
// strips: strip column bx * FWPB + (wave in block), strip row by.  Every wave
// of a block has the same rows.
__device__ __forceinline__ StagePos stage_pos_at(int bx, int by, int ny, int nx, int sr) {
    StagePos p;
    p.lane = threadIdx.x & (SW - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / SW);
    p.bx = bx * FWPB + w;
    p.live = p.bx * FO < nx;
    p.c = p.bx * FO - FHW + p.lane;
    p.cc = p.c < 0 ? 0 : (p.c >= nx ? nx - 1 : p.c);
    p.out = p.lane >= FHW && p.lane < FHW + FO && p.c < nx;
    p.hval = p.lane >= 1 && p.lane <= SW - 2 && p.c >= 0 && p.c < nx;
    p.r0 = by * sr;
    p.r1 = p.r0 + sr < ny ? p.r0 + sr : ny;
    return p;
}

__device__ __forceinline__ StagePos stage_pos(int ny, int nx, int sr) {
    int bx, by;
    xcd_block(bx, by);
    return stage_pos_at(bx, by, ny, nx, sr);
}

// Blocks of a domain's stage sweep: FWPB strip columns per block in x, one
// strip row per block in y.
__host__ __device__ inline int stage_gx(int nx) { return ((nx + FO - 1) / FO + FWPB - 1) / FWPB; }
static inline dim3 stage_grid(int ny, int nx, int sr = SR_STAGE) {
    return dim3(stage_gx(nx), (ny + sr - 1) / sr);
}

// The two-cell frame: rows [0, top) and [bot, ny), and in the rows between,
// columns [0, lft) and [rgt, nx).  Cells off the frame are "interior".
struct Frame {
    int top, bot, lft, rgt;
};

__host__ __device__ __forceinline__ Frame frame_of(int ny, int nx) {
    Frame f;
    f.top = ny < 2 ? ny : 2;
    f.bot = ny - 2 > f.top ? ny - 2 : f.top;
    f.lft = nx < 2 ? nx : 2;
    f.rgt = nx - 2 > f.lft ? nx - 2 : f.lft;
    return f;
}

__host__ __device__ __forceinline__ long frame_cells(int ny, int nx) {
    const Frame f = frame_of(ny, nx);
    return long(f.top + (ny - f.bot)) * nx + long(f.bot - f.top) * (f.lft + (nx - f.rgt));
}

// Cell (r, c) of frame index i (0 <= i < frame_cells).
__device__ __forceinline__ void frame_cell(long i, int ny, int nx, int &r, int &c) {
    const Frame f = frame_of(ny, nx);
    const long ntb = long(f.top + (ny - f.bot)) * nx;
    if (i < ntb) {
        const int k = int(i / nx);
        c = int(i - long(k) * nx);
        r = k < f.top ? k : f.bot + (k - f.top);
    } else {
        i -= ntb;
        const int per = f.lft + (nx - f.rgt);
        const int k = int(i / per), j = int(i - long(k) * per);
        r = f.top + k;
        c = j < f.lft ? j : f.rgt + (j - f.lft);
    }
}

// Raw row of the fused stage: the hydro inputs, and the bed inputs the hydro
// sweep does not read (compensated-bed low part; G in constant mode).
template <typename T>
struct SRow {
    HR<T> h;
    T zl, G;
};

// Static-field specialisations of the fused stage (template parameter ST):
// ST_NN: the grid has no nodata cell, so the mask is not loaded; ST_ZU (only
// with ST_NN): the fixed bed z holds one value, read once, so z is not
// streamed.  Each is one memory stream of the sweep fewer: 4-5% of its time at
// 8192^2 for the mask, 2-4% for z (the sweeps are held by their access
// pattern, which costs about as much per stream as per byte).
constexpr int ST_NN = 1, ST_ZU = 2;

// LO: the compensated-bed low part is present; MPM: G is computed, not read;
// ST: the static-field specialisation above; zu: the uniform z (ST_ZU).  All
// are template parameters so that the row refill has no branches inside it (a
// branch there makes the compiler copy the fresh mask byte behind a wait for
// every outstanding load).
// PK: in's h, hu, hv, z_b are the packed records at pq (PXH_PACKED_STAGE).
template <typename T, bool LO, bool MPM, int ST = 0, bool PK = false>
__device__ __forceinline__ SRow<T> sr_load(const F<T> &in, const T *zb_lo_in, const T *G_in, const uint8_t *m0,
                                           unsigned ro, unsigned c, T zu = T(0), const T *pq = nullptr) {
    SRow<T> q;
    if constexpr (PK) q.h = hr_load_q<(ST & ST_NN) != 0, (ST & ST_NN) && (ST & ST_ZU)>(in, pq, m0, ro, c, zu);
    else if constexpr ((ST & ST_NN) && (ST & ST_ZU)) q.h = hr_load_static(in, ro, c, zu);
    else if constexpr (ST & ST_NN) q.h = hr_load_no_nodata(in, ro, c);
    else q.h = hr_load(in, m0, ro, c);
    q.zl = T(0);
    if constexpr (LO) q.zl = ldr(zb_lo_in, ro, c);
    q.G = T(0);
    if constexpr (!MPM) q.G = ldr(G_in, ro, c);
    return q;
}

// Elements of a domain's seam-column buffer: stage_grid(ny, nx).x strip
// columns (the narrower fused strips; the two-sweep combine's fit in it).
static inline long seam_col_elems(int ny, int nx) { return long((nx + FO - 1) / FO) * ny * SEAM_REC; }

// MODE 0 (a stage on its own): out.{h,hu,hv} of every owned cell;
//   out.{z_b, z_b_lo (, G when write_G)} of the interior owned cells.
// MODE 1 (the second SSPRK2 stage): cb.{h,hu,hv} of the owned cells within
//   three cells of the grid edge (what the frame's bed update and the ring
//   fill read); old = 0.5 old + 0.5 result over h, hu, hv, G, z_b, z_b_lo of
//   the interior owned cells; with DT, the CFL minimum over the edges between
//   interior cells of the strip of the combined state into partial[block]
//   (the edges across strip seams are left to band_dt_cell), and the combined
//   dt inputs of the strip's first and last owned columns into xs (seam_rec).
// in: the stage input (hydro fields, z, n_b, z_b, z_b_lo, G).
// The strip p of one domain; returns this lane's dt minimum (DT).
// Everything by value: a reference to a kernel argument reaches the inlined
// body as a generic pointer, and the fields read through it land in vector
// registers.
template <typename T, bool FAST, int MODE, bool DT, bool LO, bool MPM, int ST = 0, bool PK = false>
__device__ __forceinline__ T stage_strip_run(const StagePos p, const F<T> in, const F<T> out, const F<T> cb,
                                             const F<T> old, const uint8_t *__restrict__ m0,
                                             const T *__restrict__ dtp, int ny, int nx, T dx, const SedK &sd,
                                             int write_G, T *__restrict__ xs) {
    const Frame fr = frame_of(ny, nx);
    const T zero = T(0), hf = T(0.5);
    const int c = p.c;
    T vmin = inf_v<T>();
    // MODE 1 with DT: this strip's seam records, staged in LDS row by row and
    // stored as one span after the rows (whole cache lines, not one partial
    // line per row)
    T *srec = nullptr;
    if constexpr (MODE == 1 && DT) {
        __shared__ T seam_lds[FWPB * SR_MAX * SEAM_REC];
        srec = seam_lds + __builtin_amdgcn_readfirstlane(threadIdx.x / SW) * (SR_MAX * SEAM_REC);
    }
    if (p.r0 < ny) {
        const T dt = *dtp;
        const T dtdx = dt / dx;
        const T zu = (ST & ST_ZU) ? in.z[0] : zero;  // the uniform fixed bed (ST_ZU)
        const unsigned ucc = unsigned(p.cc), uc = unsigned(c);
        const bool inE = p.lane <= SW - 3, inW = p.lane >= 2;
        const bool colI = c >= fr.lft && c < fr.rgt;           // interior column
        const bool colIE = c >= fr.lft && c + 1 < fr.rgt;      // x edge (c | c+1) between interior cells
        const bool frame3 = c <= 2 || c >= nx - 3;             // MODE 1: store the stage hydro here
        const bool eown = p.lane + 1 < FHW + FO;                // the lane right of this one is owned
        // hydro rows [hr0, hr1): the owned rows and one either side
        const int hr0 = p.r0 > 0 ? p.r0 - 1 : 0;
        const int hr1 = p.r1 + 1 < ny ? p.r1 + 1 : ny;
        const int rlast = hr1 < ny - 1 ? hr1 : ny - 1;  // last row any load reads
        // Every row access is relative to row b0, the first row the strip
        // touches, through pointers moved there: (rlast - b0 + 1) nx elements
        // fit a 32-bit byte offset (the launcher checks nx).
        const int b0 = hr0 > 0 ? hr0 - 1 : 0;
        const long sh = long(b0) * nx;
        const F<T> ins = shifted(in, sh), outs = shifted(out, sh), cbs = shifted(cb, sh), olds = shifted(old, sh);
        const uint8_t *m0s = m0 + sh;
        auto row = [&](int r) { return unsigned((r <= rlast ? r : rlast) - b0) * unsigned(nx); };
        // PK (PXH_PACKED_STAGE): the first stage's output keeps h, hu, hv, z_b
        // as one record of four per cell at out.h (MODE 0 writes the records,
        // MODE 1 reads them from in.h), moved to row b0 like the other fields
        constexpr bool PKI = PK && MODE == 1, PKO = PK && MODE == 0;
        const T *qin = nullptr;
        T *qout = nullptr;
        if constexpr (PKI) qin = in.h + 4 * sh;
        if constexpr (PKO) qout = out.h + 4 * sh;
        T ph = zero, phu = zero, phv = zero;  // PKO: the hydro of row k-1, stored with its bed
        HCarry<T> hc = hcarry_at<T, FAST, unsigned, PKI>(ins, m0s, row(hr0), row(b0), hr0 > 0, ucc, dx, qin);
        T cur_zl = zero, cur_G = zero;  // bed inputs of the current hydro row
        {
            const SRow<T> q0 = sr_load<T, LO, MPM, ST, PKI>(ins, ins.z_b_lo, ins.G, m0s, row(hr0), ucc, zu, qin);
            cur_zl = q0.zl;
            cur_G = q0.G;
        }
        // the next hydro row in flight (row k+1 at hydro row k)
        SRow<T> q1 = sr_load<T, LO, MPM, ST, PKI>(ins, ins.z_b_lo, ins.G, m0s, row(hr0 + 1), ucc, zu, qin);
        // MODE 1: the previous-step row k-1 in flight at hydro row k
        OR<T> o1 = {zero, zero, zero, zero, zero, zero};
        if (MODE == 1) o1 = or_load(olds, row(b0), ucc);
        ED<T> edp{};         // bed values of hydro row k-1
        T fn = zero;         // bed flux through the edge (k-2 | k-1)
        DD<T> dprev{};       // combined dt values of row k-2 (DT)
        bool have_dprev = false;
        wait_vm_all();
        auto body = [&](const int k) {
            // previous-step values of row k-1, consumed before their refill
            CB<T> cbn = {zero, zero, zero, zero};
            T ozb = zero, ozl = zero;
            if (MODE == 1) {
                cbn = cb_make(o1, edp.h, edp.hu, edp.hv, edp.G);
                ozb = o1.zb;
                ozl = o1.zl;
                o1 = or_load(olds, row(k), ucc);  // clamped: unconditional
            }
            const HD<T> nxt = hd_derive<T, FAST>(q1.h);
            const T nxt_zl = q1.zl, nxt_G = q1.G;
            q1 = sr_load<T, LO, MPM, ST, PKI>(ins, ins.z_b_lo, ins.G, m0s, row(k + 2), ucc, zu, qin);  // clamped
            // hydro row k
            const T zb_k = hc.cur.zb;
            const int nod_k = hc.cur.flags & 2;
            T h1, hu1, hv1;
            hydro_row<T, FAST, PKI>(hc, nxt, k, hr0, c, p.hval, inE, inW, ny, nx, dx, dt, dtdx, ins, m0s, sh, h1, hu1,
                                    hv1, qin);
            const bool own_k = p.out && k >= p.r0 && k < p.r1;
            if (own_k) {
                const unsigned ro = row(k);
                if (MODE == 0 && PKO) {
                    // stored with the row's bed on the next row (below)
                } else if (MODE == 0) {
                    str(outs.h, ro, uc, h1);
                    str(outs.hu, ro, uc, hu1);
                    str(outs.hv, ro, uc, hv1);
                } else if (frame3 || k <= 2 || k >= ny - 3) {
                    str(cbs.h, ro, uc, h1);
                    str(cbs.hu, ro, uc, hu1);
                    str(cbs.hv, ro, uc, hv1);
                }
            }
            ER<T> er;
            er.h = h1; er.hu = hu1; er.hv = hv1;
            er.zb = zb_k; er.G = cur_G; er.zl = cur_zl; er.m = nod_k;
            const ED<T> edk = ed_derive<T, FAST>(er, sd);
            cur_zl = nxt_zl;
            cur_G = nxt_G;
            if (k > hr0) {
                // bed edge (k-1 | k), and the bed of row k-1 when it is interior and owned
                T fy = zero;
                if (ed_active(edp, edk)) fy = ex_edge_d<1, T, FAST>(edp, edk);
                const int rb = k - 1;
                const bool rowI = rb >= fr.top && rb < fr.bot && rb >= p.r0 && rb < p.r1;
                T zst = edp.zb;  // PKO: the bed stored with row rb (frame cells: the frame sweep's to write)
                if (rowI) {
                    const ED<T> east = ed_shdn(edp);
                    T fx = zero;
                    if (c + 1 < nx && ed_active(edp, east)) fx = ex_edge_d<0, T, FAST>(edp, east);
                    const T fw = shup(fx);
                    T div = zero;
                    if (c > 0) div = div + (-fw);
                    if (c + 1 < nx) div = div + fx;
                    div = div + (-fn);  // rb >= 2
                    div = div + fy;     // rb + 1 < ny
                    if (p.out && colI) {
                        const unsigned i = row(rb);
                        const T inc = FAST ? dtdx * div : dt * div / dx;
                        T zb1 = edp.zb - inc, zl1 = zero;
                        if constexpr (LO) cb_sub(edp.zb, edp.zl, inc, zb1, zl1);
                        if (MODE == 0) {
                            if constexpr (PKO) zst = zb1;
                            else str(outs.z_b, i, uc, zb1);
                            if constexpr (LO) str(outs.z_b_lo, i, uc, zl1);
                            if (write_G) str(outs.G, i, uc, edp.G);
                        } else {
                            str(olds.h, i, uc, cbn.h);
                            str(olds.hu, i, uc, cbn.hu);
                            str(olds.hv, i, uc, cbn.hv);
                            str(olds.G, i, uc, cbn.G);
                            T zc = hf * ozb + hf * zb1;
                            if constexpr (LO) {
                                T zcl;
                                cb_half_sum(ozb, ozl, zb1, zl1, zc, zcl);
                                str(olds.z_b_lo, i, uc, zcl);
                            }
                            str(olds.z_b, i, uc, zc);
                        }
                    }
                    if (MODE == 1 && DT) {
                        // edges between interior cells of this strip: (c | c+1)
                        // of row rb, and (rb-1 | rb) when row rb-1 is owned too.
                        // The combined value of a neighbouring strip's cell is not
                        // in reach: that strip may already have stored it over
                        // `old`.  The frame dt kernel takes the seam edges.
                        const DD<T> dcur = dd_make<T, FAST>(cbn.h, cbn.hu, cbn.hv, cbn.G, edp.flags & 2);
                        if (p.out && colI && (p.lane == FHW || p.lane == FHW + FO - 1))
                            seam_put(srec + (rb - p.r0) * SEAM_REC, p.lane != FHW, cbn.h, cbn.hu, cbn.hv, cbn.G,
                                     (edp.flags & 2) != 0);
                        const DD<T> deast = dd_shdn(dcur);
                        if (p.out && eown && colIE && dd_active(dcur, deast))
                            vmin = pmin(vmin, dt_edge_d<0, T, FAST>(dcur, deast, dx));
                        if (p.out && colI && have_dprev && dd_active(dprev, dcur))
                            vmin = pmin(vmin, dt_edge_d<1, T, FAST>(dprev, dcur, dx));
                        dprev = dcur;
                        have_dprev = true;
                    }
                }
                fn = fy;
                if constexpr (PKO) {
                    if (p.out && rb >= p.r0 && rb < p.r1) str4(qout, row(rb), uc, f4v{ph, phu, phv, zst});
                }
            }
            if constexpr (PKO) {
                ph = h1;
                phu = hu1;
                phv = hv1;
            }
            edp = edk;
            // lockstep with the block's other strips (every wave of the block
            // runs the same rows)
            if constexpr (FWPB > 1) __syncthreads();
        };
        if constexpr (MODE == 0) {
            // two rows per trip: the rows carried from one trip to the next
            // alternate between two register sets instead of being copied
            // (about 40 moves a row); the combine sweep's larger body loses
            // by it (measured at 8192^2: stage 1 -5%, stage 2 +4%)
            int k = hr0;
            for (; k + 1 < hr1; k += 2) {
                body(k);
                body(k + 1);
            }
            if (k < hr1) body(k);
            if constexpr (PKO) {
                // the last hydro row has no row below it in this strip: owned
                // only on the grid's last row, a frame row
                const int kl = hr1 - 1;
                if (p.out && kl >= p.r0 && kl < p.r1) str4(qout, row(kl), uc, f4v{ph, phu, phv, edp.zb});
            }
        } else {
            for (int k = hr0; k < hr1; ++k) body(k);
        }
    }
    if constexpr (MODE == 1 && DT) {
        __syncthreads();  // every wave of the block reaches it
        const int rlo = p.r0 > fr.top ? p.r0 : fr.top, rhi = p.r1 < fr.bot ? p.r1 : fr.bot;
        if (p.r0 < ny && rhi > rlo && p.live) {
            // rows [rlo, rhi) hold the records of the interior owned rows; a
            // side whose column is not interior holds no record and is never read
            const int n = (rhi - rlo) * SEAM_REC;
            T *dst = seam_rec(xs, p.bx, rlo, ny);
            const T *src = srec + (rlo - p.r0) * SEAM_REC;
            for (int i = p.lane; i < n; i += SW) dst[i] = src[i];
        }
    }
    return vmin;
}

template <typename T, bool FAST, int MODE, bool DT, bool LO, bool MPM, int ST, bool PK>
__global__ __launch_bounds__(SW *FWPB) void k_rx_stage_strip(F<T> in, F<T> out, F<T> cb, F<T> old,
                                                             const uint8_t *__restrict__ m0,
                                                             const T *__restrict__ dtp, int ny, int nx, int sr,
                                                             T dx, SedK sd, int write_G,
                                                             T *__restrict__ partial, T *__restrict__ xs) {
    __shared__ T red[FWPB];
    const StagePos p = stage_pos(ny, nx, sr);
    T vmin = stage_strip_run<T, FAST, MODE, DT, LO, MPM, ST, PK>(p, in, out, cb, old, m0, dtp, ny, nx, dx, sd,
                                                                 write_G, xs);
    if (MODE == 1 && DT) {
        vmin = wave_block_min_n<FWPB>(vmin, red);
        if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = vmin;
    }
}

// Dynamic LDS that caps a sweep of FWPB-wavefront blocks at `waves`
// wavefronts per SIMD (0: no cap; the registers set the occupancy).  A CU's
// four SIMDs share its LDS, so with this allocation 4 waves / FWPB blocks fit
// on a CU and one more does not (allocations rounded to 512 bytes).  Fewer
// resident strips keep fewer row streams open at once, which the access
// pattern alone prefers (a probe of the second stage's loads and stores:
// 1297 us at 3 waves per SIMD, 1173 us at 2), against the latency hiding the
// arithmetic needs; on the 8192^2 step the registers' occupancy wins (2.12 ms
// uncapped; 2.32 / 2.48 ms with the first / second stage held one wave per
// SIMD lower).  PXH_STAGE1_WAVES / PXH_STAGE2_WAVES; pyexner_amd.autotune
// times the caps with the strip height.
static inline int lds_per_cu() {
    static int bytes = 0;  // one GPU type per process
    if (!bytes) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
            v > 0)
            bytes = v;
        else
            bytes = 65536;
    }
    return bytes;
}

template <typename K>
static size_t wave_cap_lds(K *kernel, int waves) {
    if (waves <= 0) return 0;
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void *>(kernel)) != hipSuccess) return 0;
    const long blocks = (4L * waves) / FWPB;
    const long per = blocks > 0 ? (long(lds_per_cu()) / blocks) & ~511L : 0;
    const long dyn = per - long(a.sharedSizeBytes);
    return dyn > 0 ? size_t(dyn) : 0;
}

// Host launcher: picks the (LO, MPM, ST) instance for in's compensated bed,
// the sediment mode and the grid's PXH_NO_NODATA / PXH_Z_UNIFORM promises
// (the uniform-z instances exist for float64 and the float32 hardware-math
// path; the IEEE float32 path streams z).  In a compensated run every set
// (in, out, old) has z_b_lo.  waves: the occupancy cap (wave_cap_lds).
// packed (float32 hardware math; first stage with out, second with in the
// first stage's output): that output's h, hu, hv, z_b are four-float records
// at its h (PXH_PACKED_STAGE).
template <typename T, bool FAST, int MODE, bool DT>
static void launch_stage(dim3 grid, hipStream_t st, const F<T> &in, const F<T> &out, const F<T> &cb,
                         const F<T> &old, const uint8_t *m0, const T *dtp, int ny, int nx, int sr, T dx,
                         const SedK &sd, int write_G, T *partial, T *xs, bool no_nodata, int waves = 0,
                         bool z_uniform = false, bool packed = false) {
    const bool lo = sizeof(T) == 4 && in.z_b_lo != nullptr;
    constexpr bool has_zu = FAST || sizeof(T) == 8;
    constexpr bool has_pk = FAST && sizeof(T) == 4;
#define PXH_STAGE_PK(LO_, MPM_, ST_, PK_)                                                                          \
    do {                                                                                                           \
        auto *kern = &k_rx_stage_strip<T, FAST, MODE, DT, LO_, MPM_, ST_, PK_>;                                     \
        hipLaunchKernelGGL(kern, grid, dim3(SW * FWPB), wave_cap_lds(kern, waves), st, in, out, cb, old, m0, dtp, \
                           ny, nx, sr, dx, sd, write_G, partial, xs);                                               \
    } while (0)
#define PXH_STAGE(LO_, MPM_, ST_)                      \
    do {                                               \
        if constexpr (has_pk) {                        \
            if (packed) {                              \
                PXH_STAGE_PK(LO_, MPM_, ST_, true);    \
                break;                                 \
            }                                          \
        }                                              \
        PXH_STAGE_PK(LO_, MPM_, ST_, false);           \
    } while (0)
#define PXH_STAGE_NN(LO_, MPM_)                                                 \
    do {                                                                        \
        if constexpr (has_zu) {                                                 \
            if (no_nodata && z_uniform) {                                       \
                PXH_STAGE(LO_, MPM_, ST_NN | ST_ZU);                            \
                break;                                                          \
            }                                                                   \
        }                                                                       \
        if (no_nodata) PXH_STAGE(LO_, MPM_, ST_NN);                             \
        else PXH_STAGE(LO_, MPM_, 0);                                           \
    } while (0)
    if constexpr (sizeof(T) == 4) {
        if (lo) {
            if (sd.mpm) PXH_STAGE_NN(true, true);
            else PXH_STAGE_NN(true, false);
            return;
        }
    }
    if (sd.mpm) PXH_STAGE_NN(false, true);
    else PXH_STAGE_NN(false, false);
#undef PXH_STAGE_NN
#undef PXH_STAGE
#undef PXH_STAGE_PK
}

// ---------------------------------------------------------------------------
// The frame, after the ring fill of the stage's hydro fields
// ---------------------------------------------------------------------------
// Per-lane-addressed loads for the frame kernels (any cell in any lane).
template <typename T>
__device__ __forceinline__ T ldl(const T *p, long i) {
    return ((const PXH_GLOBAL T *)p)[i];
}

template <typename T, bool FAST>
// hs / zs: the element strides of hyd's h, hu, hv and of zb_in (4 for the
// packed records of PXH_PACKED_STAGE, else 1).
__device__ __forceinline__ ED<T> ed_load_lane(const F<T> &hyd, const T *zb_in, const T *zb_lo_in, const T *G_in,
                                              const uint8_t *m0, const SedK &sd, long i, int hs = 1, int zs = 1) {
    ER<T> q;
    q.h = ldl(hyd.h, i * hs); q.hu = ldl(hyd.hu, i * hs); q.hv = ldl(hyd.hv, i * hs);
    q.zb = ldl(zb_in, i * zs);
    q.G = sd.mpm ? T(0) : ldl(G_in, i);
    q.zl = T(0);
    if constexpr (sizeof(T) == 4) {
        if (zb_lo_in) q.zl = ldl(zb_lo_in, i);
    }
    q.m = ldl(m0, i);
    return ed_derive<T, FAST>(q, sd);
}

template <typename T, bool FAST>
__device__ __forceinline__ ED<T> ed_at(const F<T> &hyd, const F<T> &bed, const uint8_t *m0, const SedK &sd,
                                       long i, int hs = 1, int zs = 1) {
    return ed_load_lane<T, FAST>(hyd, bed.z_b, bed.z_b_lo, bed.G, m0, sd, i, hs, zs);
}

// compute_G + exner_solve_2D of the frame's cells from the filled hyd and the
// stage input bed; MODE 0: out.{z_b, z_b_lo (, G)}; MODE 1: old = 0.5 old +
// 0.5 result over h, hu, hv, G, z_b, z_b_lo.
template <typename T, bool FAST, int MODE>
// hs: the element stride of hyd's h, hu, hv; zs: of bed.z_b; os: of out.z_b
// (4 where they are the packed records of PXH_PACKED_STAGE, else 1).
__device__ __forceinline__ void band_bed_cell(long t, const F<T> hyd, const F<T> bed, const F<T> out, const F<T> old,
                                              const uint8_t *__restrict__ m0, const T *__restrict__ dtp, int ny,
                                              int nx, T dx, const SedK &sd, int write_G, int hs = 1, int zs = 1,
                                              int os = 1) {
    int r, c;
    frame_cell(t, ny, nx, r, c);
    const T zero = T(0), hf = T(0.5);
    const T dt = *dtp;
    const long i = long(r) * nx + c;
    const ED<T> me = ed_at<T, FAST>(hyd, bed, m0, sd, i, hs, zs);
    const bool hasS = r + 1 < ny;
    T fx = zero, fw = zero, fn = zero, fy = zero;
    if (c + 1 < nx) {
        const ED<T> e = ed_at<T, FAST>(hyd, bed, m0, sd, i + 1, hs, zs);
        if (ed_active(me, e)) fx = ex_edge_d<0, T, FAST>(me, e);
    }
    if (c > 0) {
        const ED<T> w = ed_at<T, FAST>(hyd, bed, m0, sd, i - 1, hs, zs);
        if (ed_active(w, me)) fw = ex_edge_d<0, T, FAST>(w, me);
    }
    if (r > 0) {
        const ED<T> n = ed_at<T, FAST>(hyd, bed, m0, sd, i - nx, hs, zs);
        if (ed_active(n, me)) fn = ex_edge_d<1, T, FAST>(n, me);
    }
    if (hasS) {
        const ED<T> s = ed_at<T, FAST>(hyd, bed, m0, sd, i + nx, hs, zs);
        if (ed_active(me, s)) fy = ex_edge_d<1, T, FAST>(me, s);
    }
    T div = zero;
    if (c > 0) div = div + (-fw);
    if (c + 1 < nx) div = div + fx;
    if (r > 0) div = div + (-fn);
    if (hasS) div = div + fy;
    const T inc = FAST ? (dt / dx) * div : dt * div / dx;
    T zb1 = me.zb - inc, zl1 = zero;
    if constexpr (sizeof(T) == 4) {
        if (bed.z_b_lo) cb_sub(me.zb, me.zl, inc, zb1, zl1);
    }
    if (MODE == 0) {
        out.z_b[i * os] = zb1;
        if constexpr (sizeof(T) == 4) {
            if (out.z_b_lo) out.z_b_lo[i] = zl1;
        }
        if (write_G) out.G[i] = me.G;
    } else {
        const T ozb = old.z_b[i];
        T ozl = zero;
        if constexpr (sizeof(T) == 4) {
            if (old.z_b_lo) ozl = old.z_b_lo[i];
        }
        old.h[i] = hf * old.h[i] + hf * me.h;
        old.hu[i] = hf * old.hu[i] + hf * me.hu;
        old.hv[i] = hf * old.hv[i] + hf * me.hv;
        old.G[i] = hf * old.G[i] + hf * me.G;
        T zc = hf * ozb + hf * zb1;
        if constexpr (sizeof(T) == 4) {
            if (old.z_b_lo) {
                T zcl;
                cb_half_sum(ozb, ozl, zb1, zl1, zc, zcl);
                old.z_b_lo[i] = zcl;
            }
        }
        old.z_b[i] = zc;
    }
}

template <typename T, bool FAST, int MODE>
__global__ __launch_bounds__(256) void k_rx_band_bed(F<T> hyd, F<T> bed, F<T> out, F<T> old,
                                                      const uint8_t *__restrict__ m0, const T *__restrict__ dtp,
                                                      int ny, int nx, T dx, SedK sd, int write_G,
                                                      long nband, int hs, int zs, int os) {
    const long t = long(blockIdx.x) * 256 + threadIdx.x;
    if (t < nband) band_bed_cell<T, FAST, MODE>(t, hyd, bed, out, old, m0, dtp, ny, nx, dx, sd, write_G, hs, zs, os);
}

template <typename T, bool FAST>
__device__ __forceinline__ DD<T> dd_at(const F<T> &s, const uint8_t *m0, long i) {
    return dd_make<T, FAST>(s.h[i], s.hu[i], s.hv[i], s.G[i], m0[i]);
}

// Minimum over a 256-thread block, written by thread 0 to *slot.
template <typename T>
__device__ __forceinline__ void block256_min_to(T v, T *slot) {
    __shared__ T red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) v = pmin(v, red[k]);
        *slot = v;
    }
}

// Seam edges of a combine sweep.  The sweep writes the combined state over
// `old` in place, so a strip cannot take the combined value of a cell another
// strip owns; the edges between two strips are reduced after the sweep: the
// y edges (r-1 | r) at r = rs, 2 rs, ... < ny from the stored state (rows
// read whole), the x edges from the seam-column buffer the sweep filled
// (seam_rec), x seam k lying between strip columns k and k+1.
static __host__ __device__ inline long seam_y_edges(int ny, int nx, int rs) { return long((ny - 1) / rs) * nx; }

// The CFL bound of y seam edge u = k nx + c.
template <typename T, bool FAST>
__device__ __forceinline__ T seam_dt_y(long u, const F<T> &s, const uint8_t *__restrict__ m0, int nx, T dx,
                                       int rs) {
    const long k = u / nx;
    const int c = int(u - k * nx);
    const long i = (k + 1) * rs * long(nx) + c;
    const DD<T> n = dd_at<T, FAST>(s, m0, i - nx), me = dd_at<T, FAST>(s, m0, i);
    return dd_active(n, me) ? dt_edge_d<1, T, FAST>(n, me, dx) : inf_v<T>();
}

// The CFL bound of x seam edge u = k nrows + (r - r_lo) (rows side by side).
template <typename T, bool FAST>
__device__ __forceinline__ T seam_dt_x(long u, const T *__restrict__ xs, int ny, T dx, int r_lo, int nrows) {
    const int k = int(u / nrows);
    const int r = r_lo + int(u - long(k) * nrows);
    const T *L = seam_rec(xs, k, r, ny);      // side 1: the left strip's last column
    const T *R = seam_rec(xs, k + 1, r, ny);  // side 0: the right strip's first column
    const DD<T> w = dd_make<T, FAST>(L[4], L[5], L[6], L[7], L[9] != T(0));
    const DD<T> me = dd_make<T, FAST>(R[0], R[1], R[2], R[3], R[8] != T(0));
    return dd_active(w, me) ? dt_edge_d<0, T, FAST>(w, me, dx) : inf_v<T>();
}

// The two-sweep combine (SR_BED-row, SO_WIDE-column bed strips, every row
// combined by the strips): y seams, then x seams over rows [0, ny).
static __host__ __device__ inline long seam_x_edges_2s(int ny, int nx) { return long((nx - 1) / SO_WIDE) * ny; }
static inline long seam_edges_2s(int ny, int nx) { return seam_y_edges(ny, nx, SR_BED) + seam_x_edges_2s(ny, nx); }
static inline int seam_blocks(int ny, int nx) { return int((seam_edges_2s(ny, nx) + 255) / 256); }

template <typename T, bool FAST>
__global__ __launch_bounds__(256) void k_rx_seam_dt(F<T> s, const uint8_t *__restrict__ m0, const T *__restrict__ xs,
                                                     int ny, int nx, T dx, T *__restrict__ partial) {
    const long t = long(blockIdx.x) * 256 + threadIdx.x;
    const long ny_e = seam_y_edges(ny, nx, SR_BED);
    T v = inf_v<T>();
    if (t < ny_e) v = seam_dt_y<T, FAST>(t, s, m0, nx, dx, SR_BED);
    else if (t - ny_e < seam_x_edges_2s(ny, nx)) v = seam_dt_x<T, FAST>(t - ny_e, xs, ny, dx, 0, ny);
    block256_min_to(v, partial + blockIdx.x);
}

// The fused stage's seams (sr-row, FO-column strips): its sweep combines the
// cells off the two-cell frame only, so its x seams are those between two
// interior columns, over the interior rows; the frame's edges are the frame
// band's.
__host__ __device__ inline int fused_x_seams(int ny, int nx) {
    const Frame f = frame_of(ny, nx);
    return f.rgt > 0 ? (f.rgt - 1) / FO : 0;
}
__host__ __device__ inline long fused_seam_edges(int ny, int nx, int sr) {
    const Frame f = frame_of(ny, nx);
    return seam_y_edges(ny, nx, sr) + long(fused_x_seams(ny, nx)) * (f.bot - f.top);
}

// CFL minimum over every edge with a frame cell on either side (an edge
// between two frame cells may be evaluated twice; the minimum does not mind),
// of the combined state s, for t < nband; for nband <= t the seam edges of the
// fused stage's sr-row, FO-column strips (xs: its seam-column buffer).
template <typename T, bool FAST>
__device__ __forceinline__ T band_dt_cell(long t, long nband, long nwork, const F<T> s,
                                          const uint8_t *__restrict__ m0, const T *__restrict__ xs, int ny, int nx,
                                          T dx, int sr, unsigned shared = 0) {
    T v = inf_v<T>();
    if (t < nband) {
        int r, c;
        frame_cell(t, ny, nx, r, c);
        // shared (PXH_HALO_* bits): sides whose outer two cells are a tile's
        // ghost rings, which hold no valid combined state; edges touching
        // them are left to the neighbour that owns those cells
        auto ok = [&](int rr, int cc) {
            return !(((shared & PXH_HALO_N) && rr < 2) || ((shared & PXH_HALO_S) && rr >= ny - 2) ||
                     ((shared & PXH_HALO_W) && cc < 2) || ((shared & PXH_HALO_E) && cc >= nx - 2));
        };
        if (!ok(r, c)) return v;
        const long i = long(r) * nx + c;
        const DD<T> me = dd_at<T, FAST>(s, m0, i);
        if (c + 1 < nx && ok(r, c + 1)) {
            const DD<T> e = dd_at<T, FAST>(s, m0, i + 1);
            if (dd_active(me, e)) v = pmin(v, dt_edge_d<0, T, FAST>(me, e, dx));
        }
        if (c > 0 && ok(r, c - 1)) {
            const DD<T> w = dd_at<T, FAST>(s, m0, i - 1);
            if (dd_active(w, me)) v = pmin(v, dt_edge_d<0, T, FAST>(w, me, dx));
        }
        if (r + 1 < ny && ok(r + 1, c)) {
            const DD<T> so = dd_at<T, FAST>(s, m0, i + nx);
            if (dd_active(me, so)) v = pmin(v, dt_edge_d<1, T, FAST>(me, so, dx));
        }
        if (r > 0 && ok(r - 1, c)) {
            const DD<T> n = dd_at<T, FAST>(s, m0, i - nx);
            if (dd_active(n, me)) v = pmin(v, dt_edge_d<1, T, FAST>(n, me, dx));
        }
    } else if (t < nwork) {
        const long u = t - nband, ny_e = seam_y_edges(ny, nx, sr);
        if (u < ny_e) {
            v = seam_dt_y<T, FAST>(u, s, m0, nx, dx, sr);
        } else {
            const Frame f = frame_of(ny, nx);
            v = seam_dt_x<T, FAST>(u - ny_e, xs, ny, dx, f.top, f.bot - f.top);
        }
    }
    return v;
}

// Work items of the frame dt kernel: the frame's cells, then the fused
// stage's seam edges.
static __host__ __device__ inline long band_dt_work(int ny, int nx, int sr) {
    return frame_cells(ny, nx) + fused_seam_edges(ny, nx, sr);
}

// One partial per block of the frame and seam edges, folded with a share of
// the nprev partial minima `prev` of the stage sweep (grid-strided), so the
// final reduction reads this kernel's partials only.
template <typename T, bool FAST>
__global__ __launch_bounds__(256) void k_rx_band_dt(F<T> s, const uint8_t *__restrict__ m0, const T *__restrict__ xs,
                                                     int ny, int nx, T dx, int sr, long nband, long nwork,
                                                     const T *__restrict__ prev, int nprev, T *__restrict__ partial) {
    const long t = long(blockIdx.x) * 256 + threadIdx.x;
    T v = band_dt_cell<T, FAST>(t, nband, nwork, s, m0, xs, ny, nx, dx, sr);
    for (long k = t; k < nprev; k += long(gridDim.x) * 256) v = pmin(v, prev[k]);
    block256_min_to(v, partial + blockIdx.x);
}

}  // namespace pxh
