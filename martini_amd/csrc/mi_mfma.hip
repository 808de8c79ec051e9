// Precision form of the insertion contraction on the matrix cores:
//
//   cube[(i, j), c] += sum_p W[(i, j), p] * S[p, c]
//
// per cell (16 x 16 pixels x 64 channels, or 8 x 16 x 128 for cubes of 128
// channels or more) as a block-sparse product of a pixels x 32 W tile and a
// 32 x channels S tile per chunk of 32 spatially sorted particles.  Both operands are split into BF16 pairs x = hi + lo
// (|x - hi - lo| <= 2^-16 |x|) when they are written to LDS, and each
// 16 x 16 x 32 block takes three v_mfma_f32_16x16x32_bf16 products -- hi*hi,
// hi*lo, lo*hi -- accumulated in FP32 (the dropped lo*lo term is below 2^-16
// of the product).
//
// Wave w owns four tile rows (M = the 16 pixels of a row) and four 16-channel
// blocks (N): with 16 x 16-pixel x 64-channel cells the rows 4w..4w+3 and all
// four blocks; with 8 x 16 x 128 cells rows 4 (w % 2) .. + 3 and blocks
// 4 (w / 2) .. + 3, each pair of waves sharing its rows, which the second of
// the two to finish reading them clears (an LDS counter per row group decides).  A (row, channel-block) product is
// skipped when no footprint of the chunk touches the row or no spectral window
// touches the block; the masks are wave-uniform, so the skips are scalar
// branches.
//
// LDS operand tiles hold 32 BF16 values (64 bytes) per pixel or channel, k
// contiguous, with the four 16-byte k-groups of a row permuted by bits 2..3 of
// the row index: the 16 rows a 16-lane group reads for one k-group then fall
// in 16 distinct 16-byte bank slots.
//
// Per chunk there are two barriers: staging -> (W, S fill) -> products, and a
// third in chunks with Gaussian-kernel particles, whose row and column factors
// are tabulated before the W fill.  Each wave clears the W rows it used right
// after reading them, so the next fill finds them zero; every fill rewrites
// the S blocks that the products will read.
// At the end of an item, accumulator blocks that are all zero are not
// written: a whole cell skips their read-modify-write of the cube, and a split
// cell's partial tile receives only the 1 KiB sub-tiles written, with a mask
// (16 bits per wave) that tells the reduction which to read.
// Accumulation order is fixed: deterministic, no atomics.
#include <hip/hip_runtime.h>

#include "mi_device.h"
#include "mi_internal.h"
#include "mi_tile.h"

namespace mi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

constexpr int KROW = KC * 2;  // bytes per LDS operand row (32 BF16)

// Byte offset of element (row, k) in an operand tile:
//   row * KROW + (((k >> 3) ^ (row >> 2)) & 3) * 16 + (k & 7) * 2,
// written with the k terms folded into one constant per thread (k_bits(k)).
__device__ inline int k_bits(int k) { return (((k >> 3) & 3) << 4) | ((k & 7) << 1); }
__device__ inline int op_offset(int row, int kb) { return row * KROW + (((row << 2) & 48) ^ kb); }

// Writes x as a BF16 pair (round to nearest even) at (row, k) of the hi and lo tiles.
__device__ inline void put_split(char* hi, char* lo, int row, int k, float x) {
    const __bf16 h = (__bf16)x;
    const __bf16 l = (__bf16)(x - (float)h);
    const int o = op_offset(row, k_bits(k));
    *reinterpret_cast<__bf16*>(hi + o) = h;
    *reinterpret_cast<__bf16*>(lo + o) = l;
}

// put_split of two values of column k, converted together (one packed
// conversion each for the hi and the lo parts).
__device__ inline void put_split2(char* hi, char* lo, int k, int row1, float x1, int row2,
                                  float x2) {
    const floatx2 x = {x1, x2};
    const bf16x2 h = __builtin_convertvector(x, bf16x2);
    const floatx2 hf = __builtin_convertvector(h, floatx2);
    const bf16x2 l = __builtin_convertvector(floatx2{x1 - hf.x, x2 - hf.y}, bf16x2);
    const int kb = k_bits(k);
    const int o1 = op_offset(row1, kb), o2 = op_offset(row2, kb);
    *reinterpret_cast<__bf16*>(hi + o1) = h.x;
    *reinterpret_cast<__bf16*>(lo + o1) = l.x;
    *reinterpret_cast<__bf16*>(hi + o2) = h.y;
    *reinterpret_cast<__bf16*>(lo + o2) = l.y;
}

// The 8 BF16 values of k-group kg of `row`.
__device__ inline bf16x8 get8(const char* tile, int row, int kg) {
    return *reinterpret_cast<const bf16x8*>(tile + row * KROW + (((kg ^ (row >> 2)) & 3) << 4));
}

// put_split2 for the S fill, whose lanes l and l ^ 1 hold the same two rows of
// particles k and k ^ 1 (rows given as 16 * block + r: the block's 1 KiB step
// plus op_offset(r)).  After one lane swap (DPP quad_perm [1, 0, 3, 2]) the
// even lane writes row1 and the odd lane row2, each as the pair (k even, k odd)
// -- adjacent BF16 slots of one 4-byte word -- in both tiles: one paired
// 4-byte store per lane (the hi and lo tiles lie a fixed distance apart)
// instead of four 2-byte ones.  Every lane of the wave must call it.
__device__ inline void put_split2_pairs(char* hi, char* lo, int k, int row1, float x1, int row2,
                                        float x2) {
    constexpr int SWAP = 0xB1;  // quad_perm [1, 0, 3, 2]: lanes l and l ^ 1
    const float y1 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x1), SWAP, 0xF, 0xF, false));
    const float y2 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x2), SWAP, 0xF, 0xF, false));
    const bool even = (k & 1) == 0;
    const int row = even ? row1 : row2;
    const floatx2 x = even ? floatx2{x1, y1} : floatx2{y2, x2};  // particles k & ~1, k | 1
    const bf16x2 h = __builtin_convertvector(x, bf16x2);
    const floatx2 hf = __builtin_convertvector(h, floatx2);
    const bf16x2 l = __builtin_convertvector(floatx2{x.x - hf.x, x.y - hf.y}, bf16x2);
    const int o = (row >> 4) * (16 * KROW) + op_offset(row & 15, k_bits(k & ~1));
    *reinterpret_cast<bf16x2*>(hi + o) = h;
    *reinterpret_cast<bf16x2*>(lo + o) = l;
}

// Three waves per SIMD: the 64 accumulators then live in VGPRs and three work
// groups share a CU (LDS 48 KB each).
template <class G>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_gather_mfma(
    const mi_particle* __restrict__ P, const uint32_t* __restrict__ order,
    const Item* __restrict__ items, const uint32_t* __restrict__ item_order,
    const uint32_t* __restrict__ status, const float4* __restrict__ wtab, Grid g, mi_cube_desc d,
    float* __restrict__ cube, float* __restrict__ partials, uint64_t* __restrict__ part_masks) {
    if (status[ST_OVERFLOW]) return;
    if (blockIdx.x >= status[ST_ITEMS]) return;
    const Item it = items[item_order[blockIdx.x]];  // longest items first
    const CellPos cp = cell_pos(it.cell, g);
    const int x0 = cp.x0, y0 = cp.y0, c0 = cp.c0;

    __shared__ __attribute__((aligned(16))) char Ah[G::PX * KROW];  // W hi [pixel][k]
    __shared__ __attribute__((aligned(16))) char Al[G::PX * KROW];  // W lo
    __shared__ __attribute__((aligned(16))) char Bh[G::CB * KROW];  // S hi [channel][k]
    __shared__ __attribute__((aligned(16))) char Bl[G::CB * KROW];  // S lo
    __shared__ ChunkState cs;
    __shared__ float4 Tk[LTAB_N];  // the item's kernel table
    __shared__ GaussTables gt;     // the chunk's Gaussian-kernel factors
    // Waves: RG groups of four tile rows x CGN groups of four 16-channel blocks.
    constexpr int RG = G::TX / 4, CGN = G::CB / 64;
    static_assert(RG * CGN == NT / 64, "one (row group, channel group) per wave");
    static_assert(SUB_CH == 16, "a sub-tile is one 16 x 16 product block");
    __shared__ uint32_t done_rows[RG];  // waves done reading each row group this chunk

    const int t = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform
    const int lane = t & 63;
    const int m = lane & 15;   // A row / B column / D column of this lane
    const int kg = lane >> 4;  // K group of 8 (A, B); D row group of 4

    for (int e = t; e < G::PX * KROW / 16; e += NT) {
        reinterpret_cast<uint4*>(Ah)[e] = make_uint4(0u, 0u, 0u, 0u);
        reinterpret_cast<uint4*>(Al)[e] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (t < RG) done_rows[t] = 0u;
    const int rh = wave % RG, ch = wave / RG;  // this wave's row group and channel group

    floatx4 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) acc[r][nb] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Every wave holds the records of the next chunk and the sort indices of the
    // one after.  A chunk is staged by the wave that had the fewest (row, block)
    // products in the chunk before (it reaches the staging first); every wave
    // picks it from the same row and channel-block masks.
    const int tkind = stage_table(P, order, it.begin, wtab, Tk);
    RecRegs next{};
    uint32_t next_idx = 0;
    if (lane < min(KC, it.end - it.begin)) next = load_record(P, order[it.begin + lane]);
    if (lane < min(KC, it.end - it.begin - KC)) next_idx = order[it.begin + KC + lane];
    int stager = 0;

    for (int kb = it.begin; kb < it.end; kb += KC) {
        const int nk = min(KC, it.end - kb);
        if (wave == stager) stage_chunk(next, lane < nk, x0, y0, c0, g, d, cs);
        if (lane < min(KC, it.end - kb - KC)) next = load_record(P, next_idx);
        if (lane < min(KC, it.end - kb - 2 * KC)) next_idx = order[kb + 2 * KC + lane];
        __syncthreads();
        const uint32_t rowmask = __builtin_amdgcn_readfirstlane(cs.rowmask);
        const uint32_t chanmask = __builtin_amdgcn_readfirstlane(cs.chanmask);
        const uint32_t rows = (rowmask >> (rh * 4)) & 15u;
        const uint32_t chans = (chanmask >> (ch * 4)) & 15u;
        int fewest = 17;
#pragma unroll
        for (int w2 = 0; w2 < NT / 64; ++w2) {
            const int c = __builtin_popcount((rowmask >> ((w2 % RG) * 4)) & 15u) *
                          __builtin_popcount((chanmask >> ((w2 / RG) * 4)) & 15u);
            if (c < fewest) {
                fewest = c;
                stager = w2;
            }
        }
        if (cs.gmask != 0u) {  // the same for every thread
            fill_gauss_tables(cs, gt);
            __syncthreads();
        }
        fill_weights(cs, x0, y0, wtab, tkind, Tk, gt, d.weights,
                     [&](int k, int px, float w) { put_split(Ah, Al, px, k, w); },
                     [&](int k, int p1, float w1, int p2, float w2) {
                         put_split2(Ah, Al, k, p1, w1, p2, w2);
                     });
        fill_spectra<G::CB>(cs, nk, c0, g.nc, d,
                     [&](int k, int c1, float s1, int c2, float s2) {
                         put_split2_pairs(Bh, Bl, k, c1, s1, c2, s2);
                     },
                     chanmask);
        __syncthreads();
        if (rows != 0u) {
            // The wave's four tile rows held together: each S operand read serves all.
            bf16x8 ah[4], al[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                ah[r] = bf16x8{};
                al[r] = bf16x8{};
                if (!(rows & (1u << r)) || chans == 0u) continue;
                const int row = (rh * 4 + r) * TILE_Y + m;
                ah[r] = get8(Ah, row, kg);
                al[r] = get8(Al, row, kg);
            }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                if (!(chans & (1u << nb))) continue;
                const int col = (ch * 4 + nb) * 16 + m;
                const bf16x8 bh = get8(Bh, col, kg), bl = get8(Bl, col, kg);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (!(rows & (1u << r))) continue;
                    floatx4 c = acc[r][nb];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bh, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[r], bl, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[r], bh, c, 0, 0, 0);
                    acc[r][nb] = c;
                }
            }
            // The last of the CGN waves sharing these rows to get here clears
            // the rows' 1 KiB in both W tiles (one 16-byte store per lane
            // each): the counter's update orders it after every such wave's
            // reads of them (LDS requests of a wave are served in order; the
            // release keeps the reads before the update).
            uint32_t before = 0u;
            if (CGN > 1 && lane == 0)
                before = __hip_atomic_fetch_add(&done_rows[rh], 1u, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
            if (CGN == 1 || __builtin_amdgcn_readlane(before, 0) == CGN - 1) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (!(rows & (1u << r))) continue;
                    const int zb = (rh * 4 + r) * TILE_Y * KROW + lane * 16;
                    *reinterpret_cast<uint4*>(Ah + zb) = make_uint4(0u, 0u, 0u, 0u);
                    *reinterpret_cast<uint4*>(Al + zb) = make_uint4(0u, 0u, 0u, 0u);
                }
                if (CGN > 1 && lane == 0) done_rows[rh] = 0u;
            }
        }
    }

    // D layout of a 16x16 block: lane holds rows 4*kg..4*kg+3 (pixels j) of column m
    // (channel).  Blocks that are all zero (no product reached them) are not
    // written; a partial tile's mask records the blocks that were.
    uint32_t mine = 0u;  // bit 4 r + nb: block (r, nb) written
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int prow = rh * 4 + r;  // tile row
        const int i = x0 + prow;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const floatx4 v4 = acc[r][nb];
            if (__ballot(v4[0] != 0.f || v4[1] != 0.f || v4[2] != 0.f || v4[3] != 0.f) == 0u)
                continue;
            mine |= 1u << (4 * r + nb);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int jl = kg * 4 + q;
                const int cl = (ch * 4 + nb) * 16 + m;
                const float v = acc[r][nb][q];
                if (it.slot < 0) {
                    const int j = y0 + jl, c = c0 + cl;
                    if (i < g.nx && j < g.ny && c < g.nc) cube[((size_t)i * g.ny + j) * g.nc + c] += v;
                } else {
                    partials[(size_t)it.slot * CELL_VOXELS +
                             sub_offset(G::CB / SUB_CH, prow, ch * 4 + nb, jl) + m] = v;
                }
            }
        }
    }
    static_assert(NT / 64 == 4, "four 16-bit parts per mask");
    if (it.slot >= 0 && lane == 0)  // this wave's part: bits 16 * wave + 4 r + nb
        reinterpret_cast<uint16_t*>(part_masks + it.slot)[wave] = (uint16_t)mine;
}

hipError_t gather_mfma(const mi_particle* d_p, const Grid& g, const mi_cube_desc& d, char* ws,
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
        launch(k_gather_mfma<GeoWide>);
    else
        launch(k_gather_mfma<GeoSquare>);
    return hipGetLastError();
}

}  // namespace mi
