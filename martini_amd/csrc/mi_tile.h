// Chunk staging shared by the FP32 and MFMA tile gathers: the cell-sorted
// particle records of a work item are taken 32 at a time, their footprints
// clipped to the cell's tile, and W evaluated once per (particle, footprint
// pixel) pair.
//
// Staging is done by wave 0 with wave-wide scans (no serial loop), and the
// next chunk's records are loaded while the current chunk is processed, so
// the global-memory latency is hidden behind the W/S/accumulate work.
#pragma once

#include <hip/hip_runtime.h>

#include "mi_device.h"
#include "mi_internal.h"

namespace mi {

constexpr int KC = 32;    // particles per chunk
constexpr int NT = 256;   // threads per work group

struct ChunkState {
    mi_particle p[KC];
    int ilo[KC], ihi[KC], jlo[KC], jhi[KC];
    int off[KC + 1];    // exclusive scan of the per-particle pair counts
    uint32_t rowmask;   // tile rows (x offsets) touched by any footprint of the chunk
    uint32_t chanmask;  // 16-channel sub-blocks of the cell touched by any window
};

// Called by all 64 lanes of wave 0: lane t < KC holds particle t of the chunk
// in `p` (valid when t < nk).  Clips footprints to the tile, scans the pair
// counts and ORs the row and channel-block masks across the wave.
__device__ inline void stage_chunk(const mi_particle& p, bool valid, int x0, int y0, int c0,
                                   const Grid& g, ChunkState& cs) {
    const int lane = threadIdx.x & 63;
    int cnt = 0, ilo = 1, ihi = 0, jlo = 1, jhi = 0;
    uint32_t rows = 0, chans = 0;
    if (valid && footprint(p, g.nx, g.ny, ilo, ihi, jlo, jhi)) {
        ilo = max(ilo, x0);
        ihi = min(ihi, x0 + TILE - 1);
        jlo = max(jlo, y0);
        jhi = min(jhi, y0 + TILE - 1);
        if (ilo <= ihi && jlo <= jhi) {
            cnt = (ihi - ilo + 1) * (jhi - jlo + 1);
            rows = (2u << (ihi - x0)) - (1u << (ilo - x0));
            const int a = max(p.c_lo, c0) - c0, b = min(p.c_hi, c0 + CBLOCK - 1) - c0;
            if (a <= b) chans = (2u << (b >> 4)) - (1u << (a >> 4));
        } else {
            ilo = 1; ihi = 0; jlo = 1; jhi = 0;
        }
    } else {
        ilo = 1; ihi = 0; jlo = 1; jhi = 0;
    }
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        rows |= __shfl_xor(rows, d, 64);
        chans |= __shfl_xor(chans, d, 64);
    }
    if (lane < KC) {
        if (valid) cs.p[lane] = p;
        cs.ilo[lane] = ilo;
        cs.ihi[lane] = ihi;
        cs.jlo[lane] = jlo;
        cs.jhi[lane] = jhi;
        cs.off[lane + 1] = incl;
    }
    if (lane == 0) {
        cs.off[0] = 0;
        cs.rowmask = rows;
        cs.chanmask = chans;
    }
}

// Evaluates W for every (particle, pixel) pair of the chunk and hands each
// value to store(k, pixel-in-tile, w).
template <class Store>
__device__ inline void fill_weights(const ChunkState& cs, int x0, int y0, const mi_cube_desc& d,
                                    const float4* __restrict__ wtab, Store store) {
    const int total = cs.off[KC];
    for (int q = threadIdx.x; q < total; q += NT) {
        int lo = 0, hi = KC - 1;  // find k with off[k] <= q < off[k+1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cs.off[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        const int k = lo;
        const int width = cs.jhi[k] - cs.jlo[k] + 1;
        const int local = q - cs.off[k];
        const int a = local / width;
        const int b = local - a * width;
        const int i = cs.ilo[k] + a, j = cs.jlo[k] + b;
        const mi_particle& p = cs.p[k];
        const double dx = (double)(p.ix - i) + (double)p.fx;
        const double dy = (double)(p.iy - j) + (double)p.fy;
        const int gi = gidx_of(p.r_kind);
        store(k, (i - x0) * TILE + (j - y0),
              pixel_weight(kind_of(p.r_kind), dx, dy, p.h, d.gauss_t[gi], d.gauss_inv_norm[gi],
                           wtab));
    }
}

// Cell coordinates of a work item.
struct CellPos {
    int x0, y0, c0;
};

__device__ inline CellPos cell_pos(int cell, const Grid& g) {
    const int cb = cell % g.ncb;
    const int txy = cell / g.ncb;
    return CellPos{(txy / g.nty) * TILE, (txy % g.nty) * TILE, cb * CBLOCK};
}

}  // namespace mi
