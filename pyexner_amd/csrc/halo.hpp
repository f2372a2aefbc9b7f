// Halo strips of a subdomain tile (make_halo_exchange, roe_exner.py:719-803).
//
// A tile holds its owned block plus hw ghost rings (hw = 1, 2 or 3).  The send
// strip of a side is the hw owned rows/columns next to that side; the receive
// strip is the hw ghost rows/columns on that side.  Depth d = 0 is the
// row/column nearest the shared boundary on both tiles, so a neighbour's send
// strip is this tile's receive strip as it is.
//
// Corners travel in two phases, as in a two-pass exchange: the west/east
// strips span every row of the tile, the north/south strips every column, and
// the caller completes the west/east exchange before it packs north/south.
// The north/south strips then carry the corner blocks the west/east phase
// delivered (the reference exchanges no corners: with one ghost ring its
// stencil never reads them, but the momentum corrections next to a tile edge
// read the updated depth of a ghost cell, whose edges reach the corner cells).
//
// Buffer layout: for each field of the selected groups in PXH_G_* order
// (z_b_lo right after z_b when present), for each selected side in the order
// west, north, east, south: hw strips of the side's length, depth-major.
#pragma once
#include "common.hpp"

namespace pxh {

constexpr int HALO_MAX_FIELDS = 9;

struct HaloPlan {
    int nf;                        // fields
    long off[HALO_MAX_FIELDS + 1]; // element offset of each field's block in the buffer
    int side_on[4];                // west, north, east, south selected
    long side_off[5];              // element offset of each side inside a field's block
};

template <typename T>
struct HaloFields {
    T *p[HALO_MAX_FIELDS];
};

// Element (k, s, d, i) of the strips: the cell it packs from (send) or unpacks
// into (receive).  s: 0 west, 1 north, 2 east, 3 south.
__device__ __forceinline__ long halo_cell(int s, int d, long i, int ny, int nx, int hw, bool recv) {
    long r, c;
    if (s == 0) {        // west: column hw + d (send) / hw - 1 - d (receive), every row
        r = i;
        c = recv ? hw - 1 - d : hw + d;
    } else if (s == 2) { // east
        r = i;
        c = recv ? nx - hw + d : nx - hw - 1 - d;
    } else if (s == 1) { // north: every column
        r = recv ? hw - 1 - d : hw + d;
        c = i;
    } else {             // south
        r = recv ? ny - hw + d : ny - hw - 1 - d;
        c = i;
    }
    return r * nx + c;
}

template <typename T, bool UNPACK>
__global__ __launch_bounds__(256) void k_halo(HaloFields<T> f, HaloPlan plan, T *__restrict__ buf, int ny, int nx,
                                              int hw) {
    const long per = plan.off[1] - plan.off[0];  // elements per field
    const long total = long(plan.nf) * per;
    for (long t = long(blockIdx.x) * 256 + threadIdx.x; t < total; t += long(gridDim.x) * 256) {
        const int k = int(t / per);
        long e = t - long(k) * per;
        int s = 0;
        while (s < 3 && (!plan.side_on[s] || e >= plan.side_off[s + 1])) ++s;
        e -= plan.side_off[s];
        const long len = (s == 0 || s == 2) ? ny : nx;
        const int d = int(e / len);
        const long i = e - long(d) * len;
        const long cell = halo_cell(s, d, i, ny, nx, hw, UNPACK);
        if (UNPACK)
            f.p[k][cell] = buf[t];
        else
            buf[t] = f.p[k][cell];
    }
}

// Batched rectangle copies (pxh_copy_batch): element t belongs to the last job
// whose start is <= t (jobs with no elements share the next job's start).
template <typename T>
__global__ __launch_bounds__(256) void k_copy_batch(const pxh_copy_job *__restrict__ jobs, int n, long total) {
    for (long t = long(blockIdx.x) * 256 + threadIdx.x; t < total; t += long(gridDim.x) * 256) {
        int lo = 0, hi = n;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (jobs[mid].start <= t) lo = mid;
            else hi = mid;
        }
        const pxh_copy_job &j = jobs[lo];
        const long e = t - j.start;
        const long r = e / j.cols, c = e - r * j.cols;
        static_cast<T *>(j.dst)[r * j.dst_ld + c] = static_cast<const T *>(j.src)[r * j.src_ld + c];
    }
}

}  // namespace pxh
