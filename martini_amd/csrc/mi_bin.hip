// On-GPU binning of particles into (pixel tile x channel block) cells.
//
// Replaces the reference's per-pixel O(N) square mask
// (martini/martini.py:271-274): every particle is listed once in each cell
// its footprint and channel window touch, and the list is sorted by
// (cell, kernel group, local order) with a stable radix sort, so that each
// cell's particles form one contiguous, deterministic segment.  The MFMA form
// keeps the record order within a cell (its records are laid out by tile and
// pixel Z order, so its chunks are spatially compact); the FP32 form orders
// by a hash of the particle index (chunks spread over the tile).
// Long segments are cut into work items of at most item_length() particles;
// such cells get one partial tile per item.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "mi_device.h"
#include "mi_internal.h"
#include "mi_tile.h"

namespace mi {

Grid make_grid(const mi_cube_desc& d) {
    Grid g;
    g.nx = d.nx;
    g.ny = d.ny;
    g.nc = d.nc;
    g.wide = wide_cells(d.nc);
    g.tx_shift = g.wide ? 3 : 4;
    g.ty_shift = 4;
    g.cb_shift = g.wide ? 7 : 6;
    static_assert(GeoWide::TX == 8 && GeoSquare::TX == 16 && GeoWide::TY == 16 &&
                      GeoSquare::TY == 16 && GeoWide::CB == 128 && GeoSquare::CB == 64,
                  "the shifts above");
    g.ntx = (d.nx + (1 << g.tx_shift) - 1) >> g.tx_shift;
    g.nty = (d.ny + (1 << g.ty_shift) - 1) >> g.ty_shift;
    g.ncb = (d.nc + (1 << g.cb_shift) - 1) >> g.cb_shift;
    g.ncells = (int64_t)g.ntx * g.nty * g.ncb;
    int b = 1;
    while (((int64_t)1 << b) < g.ncells) ++b;
    g.cell_bits = b;
    return g;
}

// Count and emit passes: EMIT_T threads per block, EMIT_ROUNDS particles per
// thread.
constexpr int EMIT_T = 256;
constexpr int EMIT_ROUNDS = 4;
constexpr int64_t EMIT_SPAN = (int64_t)EMIT_T * EMIT_ROUNDS;

__host__ __device__ inline int64_t emit_blocks(int64_t n) {
    return n > 0 ? (n + EMIT_SPAN - 1) / EMIT_SPAN : 0;
}

// Lowest key bit the coarse sort orders on: the top 16 bits of the key
// (cell, kernel group and the high bits of the index hash) fit two 8-bit radix
// passes.  Within equal sorted bits the stable sort keeps emit order.
static int coarse_begin_bit(const Grid& g) {
    const int b = g.cell_bits + KEY_LOCAL_BITS - 16;
    return b < 0 ? 0 : (b > 8 ? 8 : b);
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

hipError_t layout_for(const Grid& g, int64_t n, int64_t n_pairs, Layout* L) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t at = off;
        off = align256(off + bytes);
        return at;
    };
    size_t np = (size_t)(n_pairs > 0 ? n_pairs : 1);
    const size_t nb2 = (size_t)emit_blocks(n) + 2;
    size_t nc1 = (size_t)g.ncells + 1;
    L->item_len = item_length(n_pairs);
    L->items_max = g.ncells + n_pairs / L->item_len + 1;
    L->multi_max = 2 * (n_pairs / L->item_len) + 1;
    L->counts = take(4 * nb2);
    L->offsets = take(4 * nb2);
    L->keys = take(4 * np);
    L->vals = take(4 * np);
    L->keys2 = take(4 * np);
    L->vals2 = take(4 * np);
    L->cell_start = take(4 * (size_t)g.ncells);
    L->cell_end = take(4 * (size_t)g.ncells);
    L->item_off = take(4 * nc1);
    L->multi_off = take(4 * nc1);
    L->items = take(sizeof(Item) * (size_t)L->items_max);
    L->item_keys = take(4 * (size_t)L->items_max);
    L->item_keys2 = take(4 * (size_t)L->items_max);
    L->item_idx = take(4 * (size_t)L->items_max);
    L->item_order = take(4 * (size_t)L->items_max);
    L->partials = take(sizeof(float) * (size_t)L->multi_max * CELL_VOXELS);
    L->part_masks = take(sizeof(uint64_t) * (size_t)L->multi_max);
    L->status = take(4 * 8);
    // rocPRIM temporary storage: the larger of its two uses (the pair and the item sort).
    size_t t1 = 0, t2 = 0;
    uint32_t* nul = nullptr;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, t1, nul, nul, nul, nul, np, 0,
                                             g.cell_bits + KEY_LOCAL_BITS, 0);
    if (e != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(nullptr, t2, nul, nul, nul, nul, (size_t)L->items_max, 0,
                                  ITEM_KEY_BITS, 0);
    if (e != hipSuccess) return e;
    const size_t tb = t1 > t2 ? t1 : t2;
    L->temp_bytes = tb;
    L->temp = take(tb);
    L->total = off;
    return hipSuccess;
}

// Kernel families in a cell's sort order: tabulated closed forms first (the
// item stages the first one's table in LDS), then Gaussian, Dirac, and the
// caller's weight tables.
__device__ inline uint32_t kernel_group(int kind) {
    return kind == MI_KERNEL_GAUSS ? 1u
           : kind == MI_KERNEL_DIRAC ? 2u
           : kind == MI_KERNEL_TABLE ? 3u : 0u;
}

// Whether a record with a footprint can be inserted: a known kernel kind, and
// for MI_KERNEL_TABLE weights that lie inside the descriptor's table.
__device__ inline bool record_ok(const mi_particle& p, const Grid& g, const float* uw,
                                 int64_t uw_n) {
    const int kind = kind_of(p.r_kind);
    if (kind < MI_KERNEL_TABLE) return true;
    if (kind > MI_KERNEL_TABLE || uw == nullptr) return false;
    int i_lo, i_hi, j_lo, j_hi;
    if (!footprint(p, g.nx, g.ny, i_lo, i_hi, j_lo, j_hi)) return true;
    const int64_t end = (int64_t)table_offset(p) + (int64_t)(i_hi - i_lo + 1) * (j_hi - j_lo + 1);
    return end <= uw_n;
}

// Count and emit passes: block b takes particles b * EMIT_SPAN + i * EMIT_T + t
// (rounds i = 0 .. EMIT_ROUNDS - 1, thread t), so the record loads of a round
// are contiguous.  k_count writes each block's pair total, k_scan_blocks turns
// them into block bases, and k_emit recounts its particles and scans them in
// the block: particle k's pairs start at the sum of the counts before k, as
// one global exclusive scan would place them, without a per-particle array.

// Exclusive block-wide scan (EMIT_T threads, all active); the block total in `total`.
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t* wsum, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t incl = (uint32_t)wave_scan_add((int)x);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < EMIT_T / 64; ++i) {
        const uint32_t v = wsum[i];
        base += i < w ? v : 0u;
        tot += v;
    }
    __syncthreads();  // wsum may be written again
    total = tot;
    return base + incl - x;
}

__global__ __launch_bounds__(EMIT_T) void k_count(const mi_particle* __restrict__ P, int64_t n,
                                                  Grid g, const float* __restrict__ uw,
                                                  int64_t uw_n, uint32_t* __restrict__ block_tot,
                                                  uint32_t* __restrict__ status) {
    __shared__ uint32_t wsum[EMIT_T / 64];
    const int64_t base = (int64_t)blockIdx.x * EMIT_SPAN + threadIdx.x;
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < EMIT_ROUNDS; ++i) {
        const int64_t k = base + (int64_t)i * EMIT_T;
        if (k >= n) break;
        const mi_particle p = P[k];
        int a, b, c, d, e, f;
        const uint32_t cells = cells_of(p, g, a, b, c, d, e, f);
        sum += cells;
        if (cells > 0 && !record_ok(p, g, uw, uw_n)) status[ST_BADREC] = 1u;  // same value from all
    }
    uint32_t total;
    block_excl_scan(sum, wsum, total);
    if (threadIdx.x == 0) block_tot[blockIdx.x] = total;
}

// Exclusive scan of the nb block totals in one work group (runs per thread,
// then the run totals in LDS); base[nb], base[nb + 1] hold the 64-bit grand
// total (low, high word), so a device count beyond 2^32 is still seen.
constexpr int SCAN_T = 1024;

__global__ __launch_bounds__(SCAN_T) void k_scan_blocks(const uint32_t* __restrict__ tot,
                                                        int64_t nb, uint32_t* __restrict__ base) {
    __shared__ unsigned long long s[SCAN_T];
    const int t = threadIdx.x;
    const int64_t per = (nb + SCAN_T - 1) / SCAN_T;
    const int64_t b0 = min((int64_t)t * per, nb), b1 = min(b0 + per, nb);
    unsigned long long run = 0;
    for (int64_t b = b0; b < b1; ++b) run += tot[b];
    s[t] = run;
    __syncthreads();
    for (int d = 1; d < SCAN_T; d <<= 1) {  // inclusive scan of the run totals
        const unsigned long long v = t >= d ? s[t - d] : 0ull;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    unsigned long long at = s[t] - run;
    for (int64_t b = b0; b < b1; ++b) {
        base[b] = (uint32_t)at;
        at += tot[b];
    }
    if (t == SCAN_T - 1) {
        base[nb] = (uint32_t)s[t];
        base[nb + 1] = (uint32_t)(s[t] >> 32);
    }
}

// Writes the (key, particle) pairs of each particle at its scanned offset.
// With `mix` the low 8 key bits are a hash of the particle index: the FP32
// gather's waves each own 4 tile rows, and a chunk of particles drawn from all
// over the tile gives them even shares of the products, where a spatially
// compact chunk loads one or two of them.  Without it they are zero and a
// cell's pairs keep the record order (the MFMA form, whose records are laid
// out in tile and pixel Z order).
// Entries past the device total up to the capacity get key 0xffffffff so the
// sort moves them to the end; a total above the capacity raises the overflow
// flag and writes nothing.
__global__ __launch_bounds__(EMIT_T) void k_emit(const mi_particle* __restrict__ P, int64_t n,
                                                 int64_t cap, Grid g,
                                                 const uint32_t* __restrict__ block_base,
                                                 uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals,
                                                 uint32_t* __restrict__ status, int mix) {
    __shared__ uint32_t wsum[EMIT_T / 64];
    const int64_t nb = emit_blocks(n);
    const uint64_t total = (uint64_t)block_base[nb] | ((uint64_t)block_base[nb + 1] << 32);
    const int64_t k0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k0 == 0) {
        status[ST_PAIRS] = (uint32_t)total;
        status[ST_OVERFLOW] = total > (uint64_t)cap ? 1u : 0u;
    }
    if (total > (uint64_t)cap) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)total + k0; q < cap; q += stride) {  // padding entries
        keys[q] = 0xffffffffu;
        vals[q] = 0u;
    }
    if ((int64_t)blockIdx.x >= nb) return;  // uniform
    uint32_t run = block_base[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * EMIT_SPAN + threadIdx.x;
#pragma unroll 1
    for (int i = 0; i < EMIT_ROUNDS; ++i) {
        const int64_t k = base + (int64_t)i * EMIT_T;
        mi_particle p;
        int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1, cb0 = 0, cb1 = -1;
        uint32_t cnt = 0;
        if (k < n) {
            p = P[k];
            cnt = cells_of(p, g, tx0, tx1, ty0, ty1, cb0, cb1);
        }
        uint32_t round_total;
        uint32_t at = run + block_excl_scan(cnt, wsum, round_total);
        run += round_total;
        if (cnt == 0) continue;
        const uint32_t kg = kernel_group(kind_of(p.r_kind));
        const uint32_t hash = ((uint32_t)k * 2654435761u) >> 24;
        const uint32_t local = (kg << 8) | (mix ? hash : 0u);
        for (int tx = tx0; tx <= tx1; ++tx) {
            for (int ty = ty0; ty <= ty1; ++ty) {
                for (int cb = cb0; cb <= cb1; ++cb) {
                    uint32_t cell = (uint32_t)(((int64_t)tx * g.nty + ty) * g.ncb + cb);
                    keys[at] = (cell << KEY_LOCAL_BITS) | local;
                    vals[at] = (uint32_t)k;
                    ++at;
                }
            }
        }
    }
}

// Zeroes the status words and the cell bounds (cells without particles keep
// start = end = 0).
__global__ void k_bin_init(uint32_t* __restrict__ status, uint32_t* __restrict__ start,
                           uint32_t* __restrict__ end, int64_t ncells) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 8) status[k] = 0u;
    if (k < ncells) {
        start[k] = 0u;
        end[k] = 0u;
    }
}

// Cell bounds from the sorted keys: thread t takes keys 4t .. 4t + 3 (one
// 16-byte load when the run lies inside the array, the workspace keys being
// 256-byte aligned) and the key on either side.
__global__ void k_cell_bounds(const uint32_t* __restrict__ keys, int64_t np, int64_t ncells,
                              uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    const int64_t q0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (q0 >= np) return;
    uint32_t c[6];  // cells of keys q0 - 1 .. q0 + 4 (0xffffffff outside the array)
    if (q0 + 4 <= np) {
        const uint4 v = *reinterpret_cast<const uint4*>(keys + q0);
        c[1] = v.x >> KEY_LOCAL_BITS;
        c[2] = v.y >> KEY_LOCAL_BITS;
        c[3] = v.z >> KEY_LOCAL_BITS;
        c[4] = v.w >> KEY_LOCAL_BITS;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) c[i + 1] = q0 + i < np ? keys[q0 + i] >> KEY_LOCAL_BITS : 0xffffffffu;
    }
    c[0] = q0 > 0 ? keys[q0 - 1] >> KEY_LOCAL_BITS : 0xffffffffu;
    c[5] = q0 + 4 < np ? keys[q0 + 4] >> KEY_LOCAL_BITS : 0xffffffffu;
#pragma unroll
    for (int i = 1; i <= 4; ++i) {
        if ((int64_t)c[i] >= ncells) continue;  // padding (and past the end)
        if (c[i] != c[i - 1]) start[c[i]] = (uint32_t)(q0 + i - 1);
        if (c[i] != c[i + 1]) end[c[i]] = (uint32_t)(q0 + i);
    }
}

// Work items per cell (ceil(len / item_len), none after an overflow) and
// their exclusive scans, in one work group: thread t takes a contiguous run of
// cells, the run totals are scanned in LDS, and the offsets written.
// item_off[ncells] and multi_off[ncells] hold the totals (also status words).

__global__ __launch_bounds__(SCAN_T) void k_item_scan(const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ end,
                                                      int64_t ncells, int item_len,
                                                      uint32_t* __restrict__ status,
                                                      uint32_t* __restrict__ item_off,
                                                      uint32_t* __restrict__ multi_off) {
    __shared__ uint32_t si[SCAN_T], sm[SCAN_T];
    const int t = threadIdx.x;
    const int64_t per = (ncells + SCAN_T - 1) / SCAN_T;
    const int64_t c0 = min((int64_t)t * per, ncells), c1 = min(c0 + per, ncells);
    const bool overflow = status[ST_OVERFLOW] != 0u || status[ST_BADREC] != 0u;
    const uint32_t L = (uint32_t)item_len;
    auto items_of = [&](int64_t c) {
        const uint32_t s0 = start[c], e0 = end[c];
        return (overflow || e0 <= s0) ? 0u : (e0 - s0 + L - 1) / L;
    };
    uint32_t ai = 0, am = 0;
    for (int64_t c = c0; c < c1; ++c) {
        const uint32_t ni = items_of(c);
        ai += ni;
        am += ni > 1 ? ni : 0u;
    }
    si[t] = ai;
    sm[t] = am;
    __syncthreads();
    for (int d = 1; d < SCAN_T; d <<= 1) {  // inclusive scan of the run totals
        const uint32_t vi = t >= d ? si[t - d] : 0u, vm = t >= d ? sm[t - d] : 0u;
        __syncthreads();
        si[t] += vi;
        sm[t] += vm;
        __syncthreads();
    }
    uint32_t bi = si[t] - ai, bm = sm[t] - am;
    for (int64_t c = c0; c < c1; ++c) {
        const uint32_t ni = items_of(c);
        item_off[c] = bi;
        multi_off[c] = bm;
        bi += ni;
        bm += ni > 1 ? ni : 0u;
    }
    if (t == SCAN_T - 1) {
        item_off[ncells] = si[t];
        multi_off[ncells] = sm[t];
        status[ST_ITEMS] = si[t];
        status[ST_MULTI] = sm[t];
    }
}

// One work group per cell writes the cell's items; a cell cut into several
// items gives each a partial-tile slot.
__global__ __launch_bounds__(64) void k_build_items(const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ end,
                                                    int item_len,
                                                    const uint32_t* __restrict__ item_off,
                                                    const uint32_t* __restrict__ multi_off,
                                                    Item* __restrict__ items) {
    const uint32_t c = blockIdx.x;
    const uint32_t base = item_off[c], ni = item_off[c + 1] - base;
    if (ni == 0u) return;
    const uint32_t s0 = start[c], e0 = end[c], L = (uint32_t)item_len;
    for (uint32_t q = threadIdx.x; q < ni; q += blockDim.x) {
        Item it;
        it.cell = (int32_t)c;
        it.begin = (int32_t)(s0 + q * L);
        it.end = (int32_t)min(e0, s0 + (q + 1) * L);
        it.slot = ni > 1 ? (int32_t)(multi_off[c] + q) : -1;
        items[base + q] = it;
    }
}

// Launch order of the work items: longest first (ties in cell order), so that
// the gathers' work groups, dispatched in order, take the long items early and
// the short ones fill the tail.  The key is ITEM_MAX_LEN - length; slots past
// the item count sort last.  Only the schedule changes: every cell still sums
// its items in item order, so the cube is bitwise the same in any order.
__global__ void k_item_keys(const Item* __restrict__ items, const uint32_t* __restrict__ status,
                            int64_t items_max, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ idx) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= items_max) return;
    const uint32_t n_items = status[ST_ITEMS];
    keys[j] = (uint32_t)j < n_items ? (uint32_t)(ITEM_MAX - (items[j].end - items[j].begin))
                                    : (1u << (ITEM_KEY_BITS - 1));
    idx[j] = (uint32_t)j;
}

static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t bin_particles(const mi_particle* d_p, int64_t n, int64_t n_pairs, const Grid& g,
                         const mi_cube_desc& d, char* ws, const Layout& L, bool record_order,
                         hipStream_t s) {
    uint32_t* counts = (uint32_t*)(ws + L.counts);
    uint32_t* offsets = (uint32_t*)(ws + L.offsets);
    uint32_t* keys = (uint32_t*)(ws + L.keys);
    uint32_t* vals = (uint32_t*)(ws + L.vals);
    uint32_t* keys2 = (uint32_t*)(ws + L.keys2);
    uint32_t* vals2 = (uint32_t*)(ws + L.vals2);
    uint32_t* cstart = (uint32_t*)(ws + L.cell_start);
    uint32_t* cend = (uint32_t*)(ws + L.cell_end);
    uint32_t* ioff = (uint32_t*)(ws + L.item_off);
    uint32_t* moff = (uint32_t*)(ws + L.multi_off);
    Item* items = (Item*)(ws + L.items);
    uint32_t* status = (uint32_t*)(ws + L.status);
    void* temp = ws + L.temp;
    size_t tb = L.temp_bytes;
    const int T = 256;

    hipLaunchKernelGGL(k_bin_init, dim3(blocks_for(g.ncells > 8 ? g.ncells : 8, T)), dim3(T), 0, s,
                       status, cstart, cend, g.ncells);
    const int64_t nb = emit_blocks(n);
    if (nb > 0)
        hipLaunchKernelGGL(k_count, dim3((unsigned)nb), dim3(EMIT_T), 0, s, d_p, n, g, d.weights,
                           d.weights_n, counts, status);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SCAN_T), 0, s, counts, nb, offsets);
    // the emit grid: one block per particle block, at least enough threads
    // (up to 2^24) to fill any padding entries in a few strides
    int64_t eb = nb;
    const int64_t pad_blocks = (n_pairs + EMIT_T - 1) / EMIT_T;
    if (eb < pad_blocks) eb = pad_blocks < 65536 ? pad_blocks : 65536;
    if (eb < 1) eb = 1;
    hipLaunchKernelGGL(k_emit, dim3((unsigned)eb), dim3(EMIT_T), 0, s, d_p, n, n_pairs, g, offsets,
                       keys, vals, status, record_order ? 0 : 1);
    if (n_pairs > 0) {
        // In record order only the cell and the kernel group are sorted (two
        // radix passes up to 2^14 cells); the stable sort keeps the rest.
        const int begin_bit = record_order ? KEY_LOCAL_BITS - 2 : coarse_begin_bit(g);
        hipError_t e = rocprim::radix_sort_pairs(temp, tb, keys, keys2, vals, vals2,
                                                 (size_t)n_pairs, begin_bit,
                                                 g.cell_bits + KEY_LOCAL_BITS, s);
        if (e != hipSuccess) return e;
    }
    if (n_pairs > 0)
        hipLaunchKernelGGL(k_cell_bounds, dim3(blocks_for((n_pairs + 3) / 4, T)), dim3(T), 0, s,
                           keys2, n_pairs, g.ncells, cstart, cend);
    hipLaunchKernelGGL(k_item_scan, dim3(1), dim3(SCAN_T), 0, s, cstart, cend, g.ncells,
                       L.item_len, status, ioff, moff);
    hipLaunchKernelGGL(k_build_items, dim3((unsigned)g.ncells), dim3(64), 0, s, cstart, cend,
                       L.item_len, ioff, moff, items);
    uint32_t* ikeys = (uint32_t*)(ws + L.item_keys);
    uint32_t* iidx = (uint32_t*)(ws + L.item_idx);
    hipLaunchKernelGGL(k_item_keys, dim3(blocks_for(L.items_max, T)), dim3(T), 0, s, items, status,
                       L.items_max, ikeys, iidx);
    hipError_t e = rocprim::radix_sort_pairs(temp, tb, ikeys, (uint32_t*)(ws + L.item_keys2), iidx,
                                             (uint32_t*)(ws + L.item_order), (size_t)L.items_max,
                                             0, ITEM_KEY_BITS, s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace mi
