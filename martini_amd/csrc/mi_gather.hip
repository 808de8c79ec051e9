// Pixel-tile gather: the FP32 form of the insertion contraction
//
//   cube[i, j, c] += sum_p W[(i, j), p] * S[p, c]
//
// over one cell (16 x 16 pixels x 64 channels) per work item, replacing the
// reference's per-pixel loop (martini/martini.py:243-283, :327-362).
//
// A work group of 256 threads takes the item's particles in chunks of 32:
//   1. the chunk's records go to LDS with their footprints clipped to the tile;
//   2. W[k][pixel] is evaluated once per (particle, footprint pixel) pair --
//      pairs are enumerated over the chunk so every lane does useful
//      transcendental work -- into a zeroed 32 x 256 LDS tile;
//   3. S[k][channel] (line profile times amplitude) fills a 32 x 64 LDS tile;
//   4. each thread owns one tile row (16 pixels) x 4 channels and accumulates
//      64 FMAs per particle; a wave skips particles whose footprint misses its
//      4 rows.
// Summation follows the sorted particle order: no atomics, deterministic.
#include <hip/hip_runtime.h>

#include "mi_device.h"
#include "mi_internal.h"
#include "mi_tile.h"

namespace mi {

__global__ __launch_bounds__(NT) void k_gather_plain(
    const mi_particle* __restrict__ P, const uint32_t* __restrict__ order,
    const Item* __restrict__ items, const uint32_t* __restrict__ status, Grid g, mi_cube_desc d,
    float* __restrict__ cube, float* __restrict__ partials) {
    if (status[ST_OVERFLOW]) return;
    const int item_id = blockIdx.x;
    if ((uint32_t)item_id >= status[ST_ITEMS]) return;
    const Item it = items[item_id];
    const int cb = it.cell % g.ncb;
    const int txy = it.cell / g.ncb;
    const int ty = txy % g.nty, tx = txy / g.nty;
    const int x0 = tx * TILE, y0 = ty * TILE, c0 = cb * CBLOCK;

    __shared__ float Wt[KC][TILE_PX];
    __shared__ __attribute__((aligned(16))) float St[KC][CBLOCK];
    __shared__ ChunkState cs;

    const int t = threadIdx.x;
    const int wave = t >> 6;
    const int pg = t >> 4;  // tile row (pixel offset along x)
    const int cg = t & 15;  // 4-channel group
    const int wrow0 = x0 + wave * 4, wrow1 = wrow0 + 3;

    float acc[TILE][4];
#pragma unroll
    for (int jj = 0; jj < TILE; ++jj)
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) acc[jj][cc] = 0.0f;

    for (int kb = it.begin; kb < it.end; kb += KC) {
        const int nk = min(KC, it.end - kb);
        load_chunk(P, order, kb, nk, x0, y0, c0, g, cs);
        // zero the weight tile rows of this chunk
        for (int e = t; e < nk * TILE_PX / 4; e += NT)
            reinterpret_cast<float4*>(&Wt[0][0])[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        fill_weights(cs, x0, y0, d, [&](int k, int px, float w) { Wt[k][px] = w; });
        for (int e = t; e < nk * CBLOCK; e += NT) {
            int k = e / CBLOCK, cl = e - k * CBLOCK;
            int c = c0 + cl;
            St[k][cl] = c < g.nc ? channel_flux(cs.p[k], c, d.spectrum) : 0.0f;
        }
        __syncthreads();
        for (int k = 0; k < nk; ++k) {
            if (cs.ihi[k] < wrow0 || cs.ilo[k] > wrow1) continue;
            const float4* wr = reinterpret_cast<const float4*>(&Wt[k][pg * TILE]);
            const float4 sv = *reinterpret_cast<const float4*>(&St[k][cg * 4]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 w = wr[q];
                const float wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc[q * 4 + u][0] = fmaf(wv[u], sv.x, acc[q * 4 + u][0]);
                    acc[q * 4 + u][1] = fmaf(wv[u], sv.y, acc[q * 4 + u][1]);
                    acc[q * 4 + u][2] = fmaf(wv[u], sv.z, acc[q * 4 + u][2]);
                    acc[q * 4 + u][3] = fmaf(wv[u], sv.w, acc[q * 4 + u][3]);
                }
            }
        }
        __syncthreads();
    }

    const int i = x0 + pg;
    if (it.slot < 0) {
        if (i < g.nx) {
#pragma unroll
            for (int jj = 0; jj < TILE; ++jj) {
                int j = y0 + jj;
                if (j >= g.ny) continue;
                float* dst = cube + ((size_t)i * g.ny + j) * g.nc;
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    int c = c0 + cg * 4 + cc;
                    if (c < g.nc) dst[c] += acc[jj][cc];
                }
            }
        }
    } else {
        float4* dst = reinterpret_cast<float4*>(partials + (size_t)it.slot * TILE_PX * CBLOCK);
#pragma unroll
        for (int jj = 0; jj < TILE; ++jj)
            dst[(pg * TILE + jj) * (CBLOCK / 4) + cg] =
                make_float4(acc[jj][0], acc[jj][1], acc[jj][2], acc[jj][3]);
    }
}

// Sums the partial tiles of each split cell in item order and adds them to
// the cube.
__global__ __launch_bounds__(NT) void k_reduce(const uint32_t* __restrict__ item_counts,
                                               const uint32_t* __restrict__ multi_off,
                                               const uint32_t* __restrict__ status, Grid g,
                                               const float* __restrict__ partials,
                                               float* __restrict__ cube) {
    if (status[ST_OVERFLOW]) return;
    const int cell = blockIdx.x;
    const uint32_t ni = item_counts[cell];
    if (ni <= 1) return;
    const float* base = partials + (size_t)multi_off[cell] * TILE_PX * CBLOCK;
    const int cb = cell % g.ncb;
    const int txy = cell / g.ncb;
    const int ty = txy % g.nty, tx = txy / g.nty;
    for (int v = threadIdx.x; v < TILE_PX * CBLOCK; v += NT) {
        int px = v / CBLOCK, cl = v - px * CBLOCK;
        int i = tx * TILE + px / TILE, j = ty * TILE + px % TILE, c = cb * CBLOCK + cl;
        float s = 0.0f;
        for (uint32_t q = 0; q < ni; ++q) s += base[(size_t)q * TILE_PX * CBLOCK + v];
        if (i < g.nx && j < g.ny && c < g.nc) cube[((size_t)i * g.ny + j) * g.nc + c] += s;
    }
}

__global__ void k_weights(int kind, float gt, float ginv, const float* __restrict__ dx,
                          const float* __restrict__ dy, const float* __restrict__ h,
                          float* __restrict__ out, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = pixel_weight(kind, (double)dx[k], (double)dy[k], h[k], gt, ginv);
}

__global__ void k_spectra(const mi_particle* __restrict__ P, int64_t n, int nc, int spectrum,
                          float* __restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nc) return;
    int64_t p = e / nc;
    int c = (int)(e - p * nc);
    out[e] = channel_flux(P[p], c, spectrum);
}

hipError_t gather_plain(const mi_particle* d_p, const Grid& g, const mi_cube_desc& d, char* ws,
                        const Layout& L, float* d_cube, hipStream_t s) {
    if (L.items_max <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_plain, dim3((unsigned)L.items_max), dim3(NT), 0, s, d_p,
                       (const uint32_t*)(ws + L.vals2), (const Item*)(ws + L.items),
                       (const uint32_t*)(ws + L.status), g, d, d_cube, (float*)(ws + L.partials));
    return hipGetLastError();
}

hipError_t reduce_partials(const Grid& g, char* ws, const Layout& L, float* d_cube,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)g.ncells), dim3(NT), 0, s,
                       (const uint32_t*)(ws + L.item_counts), (const uint32_t*)(ws + L.multi_off),
                       (const uint32_t*)(ws + L.status), g, (const float*)(ws + L.partials),
                       d_cube);
    return hipGetLastError();
}

hipError_t launch_weights(int kind, float gt, float ginv, const float* dx, const float* dy,
                          const float* h, float* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, kind, gt,
                       ginv, dx, dy, h, out, n);
    return hipGetLastError();
}

hipError_t launch_spectra(const mi_particle* d_p, int64_t n, const mi_cube_desc& d, float* out,
                          hipStream_t s) {
    int64_t tot = n * d.nc;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spectra, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, d_p, n,
                       d.nc, d.spectrum, out);
    return hipGetLastError();
}

}  // namespace mi
