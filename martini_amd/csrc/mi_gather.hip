// Pixel-tile gather: the FP32 form of the insertion contraction
//
//   cube[i, j, c] += sum_p W[(i, j), p] * S[p, c]
//
// over one cell (16 x 16 pixels x 64 channels, or 8 x 16 x 128 for cubes of
// 128 channels or more) per work item, replacing the
// reference's per-pixel loop (martini/martini.py:243-283, :327-362).
//
// A work group of 256 threads takes the item's particles in chunks of 32:
//   1. the chunk's records go to LDS with their footprints clipped to the tile;
//   2. W[k][pixel] is evaluated once per (particle, footprint pixel) pair --
//      pairs are enumerated over the chunk so every lane does useful
//      transcendental work -- into a zeroed 32 x pixels LDS tile;
//   3. S[k][channel] (line profile times amplitude) fills a 32 x channels LDS tile;
//   4. each thread owns one tile row (16 pixels) x 4 channels and accumulates
//      up to 64 FMAs per particle; a wave skips particles whose footprint
//      misses its rows (4 or 2), and the 4-column groups the footprint misses.
// Summation follows the sorted particle order: no atomics, deterministic.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mi_device.h"
#include "mi_internal.h"
#include "mi_tile.h"

namespace mi {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Three waves per SIMD (at most 168 VGPRs): three work groups share a CU.
template <class G>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_gather_plain(
    const mi_particle* __restrict__ P, const uint32_t* __restrict__ order,
    const Item* __restrict__ items, const uint32_t* __restrict__ item_order,
    const uint32_t* __restrict__ status, const float4* __restrict__ wtab, Grid g, mi_cube_desc d,
    float* __restrict__ cube, float* __restrict__ partials, uint64_t* __restrict__ part_masks) {
    if (status[ST_OVERFLOW]) return;
    if (blockIdx.x >= status[ST_ITEMS]) return;
    const Item it = items[item_order[blockIdx.x]];  // longest items first
    const CellPos cp = cell_pos(it.cell, g);
    const int x0 = cp.x0, y0 = cp.y0, c0 = cp.c0;

    __shared__ __attribute__((aligned(16))) float Wt[KC][G::PX];
    // rows padded to CB + 4 floats: the S fill writes a column (32
    // particles) per store, which then spreads over the LDS banks
    __shared__ __attribute__((aligned(16))) float St[KC][G::CB + 4];
    __shared__ ChunkState cs;
    __shared__ float4 Tk[LTAB_N];  // the item's kernel table
    __shared__ GaussTables gt;     // the chunk's Gaussian-kernel factors

    const int t = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform
    constexpr int CG = G::CB / 4;              // 4-channel groups of a block
    constexpr int WROWS = G::TX / (NT / 64);   // tile rows per wave
    static_assert(G::TX * CG == NT && WROWS * (NT / 64) == G::TX,
                  "a thread per (tile row, 4-channel group)");
    const int pg = t / CG;  // tile row (pixel offset along x)
    const int cg = t % CG;  // 4-channel group
    const int lane = t & 63;

    for (int e = t; e < KC * G::PX / 4; e += NT)
        reinterpret_cast<float4*>(&Wt[0][0])[e] = make_float4(0.f, 0.f, 0.f, 0.f);

    float acc[TILE_Y][4];
#pragma unroll
    for (int jj = 0; jj < TILE_Y; ++jj)
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) acc[jj][cc] = 0.0f;

    // Every wave holds the records of the next chunk and the sort indices of the
    // one after.  A chunk is staged by the wave that had the fewest particles to
    // multiply in the chunk before (it reaches the staging first); every wave
    // picks it from the same counts.
    const int tkind = stage_table(P, order, it.begin, wtab, Tk);
    RecRegs next{};
    uint32_t next_idx = 0;
    if (lane < min(KC, it.end - it.begin)) next = load_record(P, order[it.begin + lane]);
    if (lane < min(KC, it.end - it.begin - KC)) next_idx = order[it.begin + KC + lane];
    int stager = 0;

    for (int kb = it.begin; kb < it.end; kb += KC) {
        const int nk = min(KC, it.end - kb);
        if (wave == __builtin_amdgcn_readfirstlane(stager))
            stage_chunk(next, lane < nk, x0, y0, c0, g, d, cs);
        if (lane < min(KC, it.end - kb - KC)) next = load_record(P, next_idx);
        if (lane < min(KC, it.end - kb - 2 * KC)) next_idx = order[kb + 2 * KC + lane];
        __syncthreads();
        // Particles whose footprint meets each wave's rows (wave-uniform masks);
        // this wave keeps its own, and the least loaded wave stages the next chunk.
        const int klo = cs.ilo[lane & (KC - 1)], khi = cs.ihi[lane & (KC - 1)];
        const bool valid = lane < KC && klo <= khi;
        uint32_t kmask = 0;
        int fewest = KC + 1;
#pragma unroll
        for (int w2 = 0; w2 < NT / 64; ++w2) {
            const int r0 = x0 + w2 * WROWS;
            const uint32_t m = (uint32_t)__ballot(valid && khi >= r0 && klo <= r0 + WROWS - 1);
            const int c = __builtin_popcount(m);
            if (c < fewest) {
                fewest = c;
                stager = w2;
            }
            if (w2 == wave) kmask = m;
        }
        const bool hit = lane < KC && ((kmask >> lane) & 1u);
        // Lane k: the 4-column groups of the tile that particle k's footprint
        // covers (W is zero in the others).  Read here, before the next chunk's
        // staging may overwrite the chunk state.
        uint32_t qm = 0;
        if (hit) {
            const WalkRec& wk = cs.walk[lane];
            const int j0 = (wk.ij >> 16) - y0;
            qm = (2u << ((j0 + (wk.kw >> 8) - 1) >> 2)) - (1u << (j0 >> 2));
        }
        if (cs.gmask != 0u) {  // the same for every thread
            fill_gauss_tables(cs, gt);
            __syncthreads();
        }
        fill_weights(
            cs, x0, y0, wtab, tkind, Tk, gt, d.weights, [&](int k, int px, float w) { Wt[k][px] = w; },
            [&](int k, int p1, float w1, int p2, float w2) {
                Wt[k][p1] = w1;
                Wt[k][p2] = w2;
            });
        fill_spectra<G::CB>(cs, nk, c0, g.nc, d, [&](int k, int c1, float v1, int c2, float v2) {
            St[k][c1] = v1;
            St[k][c2] = v2;
        });
        __syncthreads();
        if (kmask == 0u) continue;
        while (kmask != 0u) {
            const int k = __builtin_ctz(kmask);
            kmask &= kmask - 1u;
            const uint32_t qk = __builtin_amdgcn_readlane(qm, k);  // wave-uniform
            const f32x4* wr = reinterpret_cast<const f32x4*>(&Wt[k][pg * TILE_Y]);
            const f32x4 sv = *reinterpret_cast<const f32x4*>(&St[k][cg * 4]);
            // all five LDS reads are issued together, ahead of the skips
            f32x4 wq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) wq[q] = wr[q];
            asm volatile("" ::"v"(wq[0]), "v"(wq[1]), "v"(wq[2]), "v"(wq[3]), "v"(sv));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (!((qk >> q) & 1u)) continue;
                const f32x4 w = wq[q];
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
        // Clear this wave's band of W (its rows' pixels, read only by this wave)
        // for the next fill: BAND / 4 lanes per particle, 64 / (BAND / 4) particles
        // per store.
        constexpr int BAND = WROWS * TILE_Y;
        static_assert(BAND % 4 == 0 && 64 % (BAND / 4) == 0, "whole float4 rows per store");
#pragma unroll
        for (int k = 0; k < KC; k += 64 / (BAND / 4))
            *reinterpret_cast<float4*>(
                &Wt[k + lane / (BAND / 4)][wave * BAND + (lane % (BAND / 4)) * 4]) =
                make_float4(0.f, 0.f, 0.f, 0.f);
    }

    // A thread's 16 pixels x 4 channels that are all zero (no particle reached
    // them) are not written.  The four threads of a sub-tile (one tile row x 16
    // channels) are neighbouring lanes; a partial tile's mask gets the sub-tiles
    // with a nonzero value, and each wave stores the whole bytes of the mask
    // that cover its rows (no two waves share a byte).
    bool nz = false;
#pragma unroll
    for (int jj = 0; jj < TILE_Y; ++jj)
        nz = nz || acc[jj][0] != 0.f || acc[jj][1] != 0.f || acc[jj][2] != 0.f ||
             acc[jj][3] != 0.f;
    const uint64_t any = __ballot(nz);
    const bool quad = ((any >> (lane & ~3)) & 0xFu) != 0u;  // this sub-tile has a value
    const int i = x0 + pg;
    if (it.slot < 0) {
        if (i < g.nx && nz) {
#pragma unroll
            for (int jj = 0; jj < TILE_Y; ++jj) {
                const int j = y0 + jj;
                if (j >= g.ny) continue;
                float* dst = cube + ((size_t)i * g.ny + j) * g.nc;
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    const int c = c0 + cg * 4 + cc;
                    if (c < g.nc) dst[c] += acc[jj][cc];
                }
            }
        }
    } else {
        static_assert(SUB_CH == 16, "four float4 per sub-tile row");
        if (quad) {
            float* dst = partials + (size_t)it.slot * CELL_VOXELS + (cg & 3) * 4;
#pragma unroll
            for (int jj = 0; jj < TILE_Y; ++jj)
                *reinterpret_cast<float4*>(dst + sub_offset(G::CB / SUB_CH, pg, cg >> 2, jj)) =
                    make_float4(acc[jj][0], acc[jj][1], acc[jj][2], acc[jj][3]);
        }
        // The wave's 16 sub-tiles: lanes 4 q .. 4 q + 3 hold row
        // wave * WROWS + 4 q / CG, block (4 q % CG) / 4.
        constexpr int TXS = __builtin_ctz(G::TX);
        static_assert((1 << TXS) == G::TX && CG % 4 == 0, "tile rows a power of two");
        uint64_t mine = 0u, owned = 0u;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint64_t bit = 1ull << sub_bit(TXS, wave * WROWS + 4 * q / CG, (4 * q % CG) / 4);
            owned |= bit;
            if ((any >> (4 * q)) & 0xFu) mine |= bit;
        }
        if (lane == 0) {
            unsigned char* mb = reinterpret_cast<unsigned char*>(part_masks + it.slot);
#pragma unroll
            for (int b = 0; b < 8; ++b)
                if ((owned >> (8 * b)) & 0xFFu) mb[b] = (unsigned char)(mine >> (8 * b));
        }
    }
}

// Sums the partial tiles of each split cell in item order and adds them to
// the cube.  Block (cell, part): 256 threads x 4 voxels cover four 1 KiB
// sub-tiles of the CELL_VOXELS = 16384-voxel tile, one per wave; the item loop
// is unrolled so its loads are independent.  A partial's sub-tile that its item
// did not write (its mask bit clear) counts as zero and is not read; voxels no
// item wrote keep the cube's value.
constexpr int RED_PARTS = CELL_VOXELS / (NT * 4);

__global__ __launch_bounds__(NT) void k_reduce(const uint32_t* __restrict__ item_off,
                                               const uint32_t* __restrict__ multi_off,
                                               const uint32_t* __restrict__ status, Grid g,
                                               const float* __restrict__ partials,
                                               const uint64_t* __restrict__ part_masks,
                                               float* __restrict__ cube) {
    if (status[ST_OVERFLOW]) return;
    const int cell = blockIdx.x / RED_PARTS;
    const int part = blockIdx.x - cell * RED_PARTS;
    const uint32_t ni = item_off[cell + 1] - item_off[cell];
    if (ni <= 1) return;
    const size_t tile = (size_t)CELL_VOXELS / 4;  // float4 per partial tile
    const float4* base = reinterpret_cast<const float4*>(partials) + (size_t)multi_off[cell] * tile;
    const uint64_t* mk = part_masks + multi_off[cell];
    const int v4 = part * NT + threadIdx.x;  // float4 index within the tile
    // v4 = (sub-tile * TILE_Y + pixel) * 4 + float4 within the 16 channels
    static_assert(TILE_Y == 16 && SUB_CH == 16 && NT % 64 == 0, "one sub-tile per wave");
    const int sub = v4 >> 6;  // row * blocks + block, the same in a wave
    const int nblk = (1 << g.cb_shift) / SUB_CH;
    const int prow = sub / nblk, jl = (v4 >> 2) & 15;
    const int cl = (sub - prow * nblk) * SUB_CH + (v4 & 3) * 4;
    const int bit = sub_bit(g.tx_shift, prow, sub - prow * nblk);
    // An unwritten sub-tile's load is sent to the first 1 KiB of the partials
    // (the same for every wave, so cached) and its bits cleared: no branch, and
    // the four loads of a step stay independent.
    const uint4* src = reinterpret_cast<const uint4*>(base);
    const uint4* hot = reinterpret_cast<const uint4*>(partials) + (v4 & 63);
    auto take = [&](uint32_t q, uint64_t m) {
        const uint4 r = m ? src[q * tile + v4] : *hot;
        const uint32_t keep = 0u - (uint32_t)m;
        return make_float4(__uint_as_float(r.x & keep), __uint_as_float(r.y & keep),
                           __uint_as_float(r.z & keep), __uint_as_float(r.w & keep));
    };
    float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f);
    uint64_t seen = 0u;
    uint32_t q = 0;
    for (; q + 4 <= ni; q += 4) {
        const uint64_t m0 = mk[q + 0] >> bit & 1u, m1 = mk[q + 1] >> bit & 1u;
        const uint64_t m2 = mk[q + 2] >> bit & 1u, m3 = mk[q + 3] >> bit & 1u;
        const float4 a = take(q + 0, m0), b = take(q + 1, m1);
        const float4 c = take(q + 2, m2), e = take(q + 3, m3);
        s0.x += a.x; s0.y += a.y; s0.z += a.z; s0.w += a.w;
        s0.x += b.x; s0.y += b.y; s0.z += b.z; s0.w += b.w;
        s0.x += c.x; s0.y += c.y; s0.z += c.z; s0.w += c.w;
        s0.x += e.x; s0.y += e.y; s0.z += e.z; s0.w += e.w;
        seen |= m0 | m1 | m2 | m3;
    }
    for (; q < ni; ++q) {
        const uint64_t m0 = mk[q] >> bit & 1u;
        const float4 a = take(q, m0);
        s0.x += a.x; s0.y += a.y; s0.z += a.z; s0.w += a.w;
        seen |= m0;
    }
    if (!seen) return;
    const CellPos cp = cell_pos(cell, g);
    const int i = cp.x0 + prow, j = cp.y0 + jl;
    if (i >= g.nx || j >= g.ny) return;
    float* dst = cube + ((size_t)i * g.ny + j) * g.nc + cp.c0 + cl;
    const float sv[4] = {s0.x, s0.y, s0.z, s0.w};
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (cp.c0 + cl + u < g.nc) dst[u] += sv[u];
}

// The weight tables are constant data (the FP64 closed forms of the four
// tabulated kernels), so each device holds one copy, filled on first use.
__device__ float4 g_wtab[WTAB_KERNELS * WTAB_N];

__global__ void k_fill_wtab(float4* __restrict__ tab) { fill_wtab(tab); }

hipError_t weight_table(hipStream_t s, const float4** out) {
    static std::mutex mu;
    static const float4* table[MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (table[dev] == nullptr) {
        // First use on this device: fill on the caller's stream and wait for it,
        // so that work on any stream may read the table afterwards.
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_wtab));
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_fill_wtab, dim3((WTAB_KERNELS * WTAB_N + 255) / 256), dim3(256), 0, s,
                           static_cast<float4*>(p));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        table[dev] = static_cast<const float4*>(p);
    }
    *out = table[dev];
    return hipSuccess;
}

__global__ void k_weights(int kind, float gt, float ginv, const float* __restrict__ dx,
                          const float* __restrict__ dy, const float* __restrict__ h,
                          const float4* __restrict__ wtab, float* __restrict__ out, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = pixel_weight(kind, (double)dx[k], (double)dy[k], h[k], gt, ginv, wtab);
}

__global__ void k_spectra(const mi_particle* __restrict__ P, int64_t n, mi_cube_desc d,
                          float* __restrict__ out) {
    const int nc = d.nc;
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nc) return;
    int64_t p = e / nc;
    int c = (int)(e - p * nc);
    out[e] = channel_flux(P[p], c, d);
}

hipError_t gather_plain(const mi_particle* d_p, const Grid& g, const mi_cube_desc& d, char* ws,
                        const Layout& L, float* d_cube, hipStream_t s) {
    if (L.items_max <= 0) return hipSuccess;
    const float4* wtab = nullptr;
    hipError_t e = weight_table(s, &wtab);
    if (e != hipSuccess) return e;
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3((unsigned)L.items_max), dim3(NT), 0, s, d_p,
                           (const uint32_t*)(ws + L.vals2), (const Item*)(ws + L.items),
                           (const uint32_t*)(ws + L.item_order), (const uint32_t*)(ws + L.status),
                           wtab, g, d, d_cube, (float*)(ws + L.partials),
                           (uint64_t*)(ws + L.part_masks));
    };
    if (g.wide)
        launch(k_gather_plain<GeoWide>);
    else
        launch(k_gather_plain<GeoSquare>);
    return hipGetLastError();
}

hipError_t reduce_partials(const Grid& g, char* ws, const Layout& L, float* d_cube,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)(g.ncells * RED_PARTS)), dim3(NT), 0, s,
                       (const uint32_t*)(ws + L.item_off), (const uint32_t*)(ws + L.multi_off),
                       (const uint32_t*)(ws + L.status), g, (const float*)(ws + L.partials),
                       (const uint64_t*)(ws + L.part_masks), d_cube);
    return hipGetLastError();
}

hipError_t launch_weights(int kind, float gt, float ginv, const float* dx, const float* dy,
                          const float* h, float* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float4* wtab = nullptr;
    hipError_t e = weight_table(s, &wtab);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, kind, gt,
                       ginv, dx, dy, h, wtab, out, n);
    return hipGetLastError();
}

hipError_t launch_spectra(const mi_particle* d_p, int64_t n, const mi_cube_desc& d, float* out,
                          hipStream_t s) {
    int64_t tot = n * d.nc;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spectra, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, d_p, n, d,
                       out);
    return hipGetLastError();
}

}  // namespace mi
