// Precision form of the insertion contraction on the matrix cores:
//
//   cube[(i, j), c] += sum_p W[(i, j), p] * S[p, c]
//
// per cell (16 x 16 pixels x 64 channels) as a block-sparse product of a
// 256 x 32 W tile and a 32 x 64 S tile per chunk of 32 spatially sorted
// particles.  Both operands are split into BF16 pairs x = hi + lo
// (|x - hi - lo| <= 2^-16 |x|) and each 16 x 16 x 32 block takes three
// v_mfma_f32_16x16x32_bf16 products -- hi*hi, hi*lo, lo*hi -- accumulated in
// FP32 (the dropped lo*lo term is below 2^-16 of the product).
//
// Wave w owns tile rows 4w..4w+3 (M = the 16 pixels of a row) and all four
// 16-channel blocks (N).  A (row, channel-block) product is skipped when no
// footprint of the chunk touches the row or no spectral window touches the
// block: the sparsity is skipped at block granularity, uniformly per wave.
// W and S come from the same staging as the FP32 gather (mi_tile.h), laid
// out k-contiguous in LDS so each lane loads its 8 K values in two 16-byte
// reads.  Accumulation order is fixed: deterministic, no atomics.
#include <hip/hip_runtime.h>

#include "mi_device.h"
#include "mi_internal.h"
#include "mi_tile.h"

namespace mi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short shortx8 __attribute__((ext_vector_type(8)));

constexpr int KPAD = KC + 4;  // row stride (floats) of the k-contiguous LDS tiles

// Round-to-nearest-even BF16 bits of a finite float.
__device__ inline uint32_t bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// Splits 8 consecutive floats into BF16 hi and lo parts.
__device__ inline void split8(const float* src, shortx8& hi, shortx8& lo) {
    const float4 a = *reinterpret_cast<const float4*>(src);
    const float4 b = *reinterpret_cast<const float4*>(src + 4);
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        uint32_t h = bf16_bits(v[q]);
        float r = v[q] - __uint_as_float(h << 16);
        hi[q] = (short)h;
        lo[q] = (short)bf16_bits(r);
    }
}

__device__ inline floatx4 mfma3(const shortx8& ah, const shortx8& al, const shortx8& bh,
                                const shortx8& bl, floatx4 c) {
    const bf16x8 Ah = __builtin_bit_cast(bf16x8, ah), Al = __builtin_bit_cast(bf16x8, al);
    const bf16x8 Bh = __builtin_bit_cast(bf16x8, bh), Bl = __builtin_bit_cast(bf16x8, bl);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, c, 0, 0, 0);
    return c;
}

__global__ __launch_bounds__(NT) void k_gather_mfma(
    const mi_particle* __restrict__ sorted, const Item* __restrict__ items,
    const uint32_t* __restrict__ status, const float4* __restrict__ wtab, Grid g, mi_cube_desc d,
    float* __restrict__ cube, float* __restrict__ partials) {
    if (status[ST_OVERFLOW]) return;
    const int item_id = blockIdx.x;
    if ((uint32_t)item_id >= status[ST_ITEMS]) return;
    const Item it = items[item_id];
    const CellPos cp = cell_pos(it.cell, g);
    const int x0 = cp.x0, y0 = cp.y0, c0 = cp.c0;

    __shared__ __attribute__((aligned(16))) float Wp[TILE_PX][KPAD];  // [pixel][k]
    __shared__ __attribute__((aligned(16))) float Sp[CBLOCK][KPAD];   // [channel][k]
    __shared__ ChunkState cs;

    const int t = threadIdx.x;
    const int wave = t >> 6;
    const int lane = t & 63;
    const int m = lane & 15;  // A row / B column / D column of this lane
    const int kg = lane >> 4; // K group of 8 (A, B); D row group of 4

    floatx4 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) acc[r][nb] = floatx4{0.f, 0.f, 0.f, 0.f};

    mi_particle next;  // wave 0: the record this lane stages next
    if (wave == 0 && t < min(KC, it.end - it.begin)) next = sorted[it.begin + t];

    for (int kb = it.begin; kb < it.end; kb += KC) {
        const int nk = min(KC, it.end - kb);
        if (wave == 0) {
            stage_chunk(next, t < nk, x0, y0, c0, g, cs);
            const int nk2 = min(KC, it.end - kb - KC);
            if (t < nk2) next = sorted[kb + KC + t];
        }
        for (int e = t; e < TILE_PX * KPAD / 4; e += NT)
            reinterpret_cast<float4*>(&Wp[0][0])[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        fill_weights(cs, x0, y0, d, wtab, [&](int k, int px, float w) { Wp[px][k] = w; });
        for (int e = t; e < CBLOCK * KC; e += NT) {
            int cl = e / KC, k = e - cl * KC;
            int c = c0 + cl;
            Sp[cl][k] = (k < nk && c < g.nc) ? channel_flux(cs.p[k], c, d.spectrum) : 0.0f;
        }
        __syncthreads();
        const uint32_t rows = cs.rowmask, chans = cs.chanmask;
        shortx8 bh[4], bl[4];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
            if ((chans >> nb) & 1u) split8(&Sp[nb * 16 + m][kg * 8], bh[nb], bl[nb]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = wave * 4 + r;
            if (!((rows >> row) & 1u)) continue;
            shortx8 ah, al;
            split8(&Wp[row * TILE + m][kg * 8], ah, al);
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                if ((chans >> nb) & 1u) acc[r][nb] = mfma3(ah, al, bh[nb], bl[nb], acc[r][nb]);
        }
        __syncthreads();
    }

    // D layout of a 16x16 block: lane holds rows 4*kg..4*kg+3 (pixels j) of column m (channel).
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int prow = wave * 4 + r;  // tile row
        const int i = x0 + prow;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int jl = kg * 4 + q;
                const int cl = nb * 16 + m;
                const float v = acc[r][nb][q];
                if (it.slot < 0) {
                    const int j = y0 + jl, c = c0 + cl;
                    if (i < g.nx && j < g.ny && c < g.nc) cube[((size_t)i * g.ny + j) * g.nc + c] += v;
                } else {
                    partials[(size_t)it.slot * TILE_PX * CBLOCK + (prow * TILE + jl) * CBLOCK + cl] = v;
                }
            }
        }
    }
}

hipError_t gather_mfma(const mi_particle* d_p, const Grid& g, const mi_cube_desc& d, char* ws,
                       const Layout& L, float* d_cube, hipStream_t s) {
    if (L.items_max <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_mfma, dim3((unsigned)L.items_max), dim3(NT), 0, s,
                       (const mi_particle*)(ws + L.sorted), (const Item*)(ws + L.items),
                       (const uint32_t*)(ws + L.status), (const float4*)(ws + L.wtab), g, d,
                       d_cube, (float*)(ws + L.partials));
    return hipGetLastError();
}

}  // namespace mi
