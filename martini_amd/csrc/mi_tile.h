// Chunk staging shared by the FP32 and MFMA tile gathers: particle records
// of a work item are taken 32 at a time, their footprints clipped to the
// cell's tile, and W evaluated once per (particle, footprint pixel) pair.
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
    int off[KC + 1];
    uint32_t rowmask;   // tile rows (x offsets) touched by any footprint of the chunk
    uint32_t chanmask;  // 16-channel sub-blocks of the cell touched by any window
};

// Loads the chunk's records, clips footprints to the tile and scans the pair
// counts.  Returns the number of (particle, pixel) pairs via cs.off[nk].
__device__ inline void load_chunk(const mi_particle* __restrict__ P,
                                  const uint32_t* __restrict__ order, int kb, int nk, int x0,
                                  int y0, int c0, const Grid& g, ChunkState& cs) {
    int t = threadIdx.x;
    if (t < KC) {
        int cnt = 0;
        int ilo = 1, ihi = 0, jlo = 1, jhi = 0;
        if (t < nk) {
            mi_particle p = P[order[kb + t]];
            cs.p[t] = p;
            if (footprint(p, g.nx, g.ny, ilo, ihi, jlo, jhi)) {
                ilo = max(ilo, x0);
                ihi = min(ihi, x0 + TILE - 1);
                jlo = max(jlo, y0);
                jhi = min(jhi, y0 + TILE - 1);
                if (ilo <= ihi && jlo <= jhi) cnt = (ihi - ilo + 1) * (jhi - jlo + 1);
                else { ilo = 1; ihi = 0; jlo = 1; jhi = 0; }
            } else {
                ilo = 1; ihi = 0; jlo = 1; jhi = 0;
            }
        }
        cs.ilo[t] = ilo;
        cs.ihi[t] = ihi;
        cs.jlo[t] = jlo;
        cs.jhi[t] = jhi;
        cs.off[t + 1] = cnt;
    }
    __syncthreads();
    if (t == 0) {
        cs.off[0] = 0;
        uint32_t rows = 0, chans = 0;
        for (int k = 0; k < KC; ++k) {
            if (cs.off[k + 1] > 0) {
                for (int i = cs.ilo[k]; i <= cs.ihi[k]; ++i) rows |= 1u << (i - x0);
                int a = max(cs.p[k].c_lo, c0) - c0, b = min(cs.p[k].c_hi, c0 + CBLOCK - 1) - c0;
                for (int q = a >> 4; q <= (b >> 4); ++q) chans |= (a <= b) ? (1u << q) : 0u;
            }
            cs.off[k + 1] += cs.off[k];
        }
        cs.rowmask = rows;
        cs.chanmask = chans;
    }
    __syncthreads();
}

// Evaluates W for every (particle, pixel) pair of the chunk and hands each
// value to store(k, pixel-in-tile, w).
template <class Store>
__device__ inline void fill_weights(const ChunkState& cs, int x0, int y0, const mi_cube_desc& d,
                                    Store store) {
    int total = cs.off[KC];
    for (int q = threadIdx.x; q < total; q += NT) {
        int lo = 0, hi = KC - 1;  // find k with off[k] <= q < off[k+1]
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (cs.off[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        int k = lo;
        int width = cs.jhi[k] - cs.jlo[k] + 1;
        int local = q - cs.off[k];
        int a = local / width;
        int b = local - a * width;
        int i = cs.ilo[k] + a, j = cs.jlo[k] + b;
        const mi_particle& p = cs.p[k];
        double dx = (double)(p.ix - i) + (double)p.fx;
        double dy = (double)(p.iy - j) + (double)p.fy;
        int gi = gidx_of(p.r_kind);
        store(k, (i - x0) * TILE + (j - y0),
              pixel_weight(kind_of(p.r_kind), dx, dy, p.h, d.gauss_t[gi], d.gauss_inv_norm[gi]));
    }
}

}  // namespace mi
