// Chunk staging shared by the FP32 and MFMA tile gathers: the cell-sorted
// particle records of a work item are taken 32 at a time, their footprints
// clipped to the cell's tile, and W evaluated once per (particle, footprint
// pixel) pair.
//
// Staging is done by one wave with wave-wide DPP scans (no serial loop): the
// gathers pick the wave with the least product work in the previous chunk.
// Every wave loads the next chunk's records while the current one is
// processed, so the global-memory latency is hidden behind the W/S/accumulate
// work.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mi_device.h"
#include "mi_internal.h"

namespace mi {

constexpr int KC = 32;    // particles per chunk
constexpr int NT = 256;   // threads per work group
// An item's kernel table in LDS: the WTAB_N intervals and one zero entry
// after them, which the branch-free walk reads for q >= 1.
constexpr int LTAB_N = WTAB_N + 1;

// What the W walk needs of a particle, read with two 16-byte loads whenever
// a thread moves on to the next particle.
// For MI_KERNEL_TABLE particles px and py hold instead, as int64 bits
// (table_walk), the index of pixel (0, 0) in the caller's weight table and the
// row stride of the particle's weights.
struct __attribute__((aligned(16))) WalkRec {
    double px, py;  // particle position ix + fx, iy + fy (exact in FP64)
    double fa;      // WFactors::a: 1/h^2, or 1/(h sig)^2 for the Gaussian
    int32_t ij;     // first row | first column << 16 of the clipped footprint
    int32_t kw;     // kernel kind | footprint width << 8
};

// Weight-table walk of an MI_KERNEL_TABLE particle: weights[base + i * stride + j]
// is its weight at pixel (i, j).
struct TableWalk {
    int64_t base, stride;
};

__device__ inline TableWalk table_walk(const WalkRec& w) {
    return TableWalk{__double_as_longlong(w.px), __double_as_longlong(w.py)};
}

// What the Gaussian walk needs beyond the WalkRec: WFactors::b, t, n and e,
// at the WalkRec's 32-byte stride.
struct __attribute__((aligned(16))) GaussRec {
    float4 f;  // b, t, n, 0
    double e;
    double unused;
};

struct ChunkState {
    mi_particle p[KC];
    WalkRec walk[KC];
    GaussRec gr[KC];    // Gaussian kernel factors
    int ilo[KC], ihi[KC];
    int off[KC + 1];    // exclusive scan of the per-particle pair counts
    // exclusive scan of the threads given to each particle's pairs (16-byte
    // aligned: fill_weights reads it eight entries at a time)
    __attribute__((aligned(16))) int toff[KC + 1];
    int per;            // pairs per thread
    uint32_t rowmask;   // tile rows (x offsets) touched by any footprint of the chunk
    uint32_t chanmask;  // 16-channel sub-blocks of the cell touched by any window
    uint32_t gmask;     // particles of the Gaussian kernel with a footprint in the tile
};

// The separable factors of the Gaussian kernel's pixel weight for the chunk's
// Gaussian particles, per footprint row and column: ex[k][r] = erf span of row
// r of particle k's clipped footprint, ey[k][c] that of column c (weight_at's
// ex and ey).  Filled once per chunk, they leave one erf and the FP64
// truncation term per pixel.
struct GaussTables {
    float ex[KC][TILE_MAX];
    float ey[KC][TILE_MAX];
};

// Wave-wide inclusive scans with DPP (row shifts within 16-lane rows, then the
// row broadcasts of lanes 15 and 31): VALU moves only, no LDS round trip as
// with shuffles.  All 64 lanes must be active.
template <class Op>
__device__ inline uint32_t wave_scan(uint32_t x, Op op) {
    // dpp_ctrl: row_shr:n = 0x110 + n, row_bcast:15 = 0x142, row_bcast:31 = 0x143;
    // lanes whose source is out of range keep `old` (the identity 0 for + and |)
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));
    return x;
}

__device__ inline int wave_scan_add(int x) {
    return (int)wave_scan((uint32_t)x, [](uint32_t a, uint32_t b) { return a + b; });
}

__device__ inline uint32_t wave_or(uint32_t x) {
    return __builtin_amdgcn_readlane(wave_scan(x, [](uint32_t a, uint32_t b) { return a | b; }),
                                     63);
}

// A particle record held in registers between its prefetch and its staging,
// as three 16-byte pieces.  (A prefetched mi_particle copied as a whole struct
// can be given a per-thread LDS or scratch slot by the compiler, and the
// store into that slot then waits for the load: the prefetch would stall.)
struct RecRegs {
    uint4 a, b, c;
};

__device__ inline RecRegs load_record(const mi_particle* __restrict__ P, uint32_t i) {
    const uint4* q = reinterpret_cast<const uint4*>(P + i);
    return RecRegs{q[0], q[1], q[2]};
}

__device__ inline mi_particle unpack_record(const RecRegs& r) {
    mi_particle p;
    p.ix = (int32_t)r.a.x;
    p.iy = (int32_t)r.a.y;
    p.fx = __uint_as_float(r.a.z);
    p.fy = __uint_as_float(r.a.w);
    p.h = __uint_as_float(r.b.x);
    p.r_kind = (int32_t)r.b.y;
    p.c_lo = (int32_t)r.b.z;
    p.c_hi = (int32_t)r.b.w;
    p.k0 = (int32_t)r.c.x;
    p.fu = __uint_as_float(r.c.y);
    p.s = __uint_as_float(r.c.z);
    p.amp = __uint_as_float(r.c.w);
    return p;
}

// Entry i (0..3) of a 4-entry kernel-argument array for a lane-varying i, by
// selects over wave-uniform values: an indexed read of the kernel arguments
// would be a memory load on the staging path.
__device__ inline float pick4(const float (&v)[4], int i) {
    const float v0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v[0])));
    const float v1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v[1])));
    const float v2 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v[2])));
    const float v3 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v[3])));
    return i == 0 ? v0 : i == 1 ? v1 : i == 2 ? v2 : v3;
}

// Called by all 64 lanes of one wave: lane t < KC holds particle t of the chunk
// in `r` (valid when t < nk).  Clips footprints to the tile, scans the pair
// counts, ORs the row and channel-block masks across the wave and forms each
// particle's weight factors.
__device__ inline void stage_chunk(const RecRegs& r, bool valid, int x0, int y0, int c0,
                                   const Grid& g, const mi_cube_desc& d, ChunkState& cs) {
    const int lane = threadIdx.x & 63;
    const mi_particle p = unpack_record(r);
    // Straight-line code with selects: every lane forms the clipped footprint,
    // lanes without a particle or a footprint in the tile end with cnt = 0.
    int fi_lo = 0, fi_hi = -1, fj_lo = 0, fj_hi = -1;
    const bool fp = footprint(p, g.nx, g.ny, fi_lo, fi_hi, fj_lo, fj_hi);
    const int ci_lo = max(fi_lo, x0), ci_hi = min(fi_hi, x0 + (1 << g.tx_shift) - 1);
    const int cj_lo = max(fj_lo, y0), cj_hi = min(fj_hi, y0 + TILE_Y - 1);
    const bool ok = valid && fp && ci_lo <= ci_hi && cj_lo <= cj_hi;
    const int ilo = ok ? ci_lo : 1, ihi = ok ? ci_hi : 0;
    const int jlo = ok ? cj_lo : 1, jhi = ok ? cj_hi : 0;
    const int cnt = ok ? (ihi - ilo + 1) * (jhi - jlo + 1) : 0;
    const uint32_t rows = ok ? (2u << (ihi - x0)) - (1u << (ilo - x0)) : 0u;
    const int ca = max(p.c_lo, c0) - c0, cb = min(p.c_hi, c0 + (1 << g.cb_shift) - 1) - c0;
    const uint32_t chans = ok && ca <= cb ? (2u << (cb >> 4)) - (1u << (ca >> 4)) : 0u;
    const int incl = wave_scan_add(cnt);
    static_assert(TILE_MAX <= 16 && GeoWide::CB / 16 <= 16, "rows and channel blocks share one word");
    const uint32_t masks = wave_or(rows | (chans << 16));  // one reduction for both
    const uint32_t gauss = (uint32_t)__ballot(lane < KC && cnt > 0 &&
                                              kind_of(p.r_kind) == MI_KERNEL_GAUSS);
    // Each particle's pairs go to ceil(cnt / per) threads of their own, per =
    // ceil(total / (NT - KC)): at most total / per + KC <= NT threads in all.
    const int total = __builtin_amdgcn_readlane(incl, 63);
    const int per = max(1, (total + (NT - KC) - 1) / (NT - KC));
    // ceil(cnt / per) = floor((cnt + per - 1 + 1/2) / per): the quotient lies at
    // least 1/74 from an integer (per <= 37), far beyond the reciprocal's error
    const float rper = __builtin_amdgcn_rcpf((float)per);
    const int tincl = wave_scan_add((int)(((float)(cnt + per - 1) + 0.5f) * rper));
    if (lane < KC) {
        if (valid) cs.p[lane] = p;
        cs.ilo[lane] = ilo;
        cs.ihi[lane] = ihi;
        const int width = max(jhi - jlo + 1, 1);
        cs.off[lane + 1] = incl;
        cs.toff[lane + 1] = tincl;
        if (cnt > 0) {
            const int kind = kind_of(p.r_kind);
            WFactors f;
            if (kind == MI_KERNEL_GAUSS) {  // the truncation and norm only for the Gaussian
                const int gi = gidx_of(p.r_kind);
                f = weight_factors(kind, p.h, pick4(d.gauss_t, gi), pick4(d.gauss_inv_norm, gi));
            } else {
                f = weight_factors(kind, p.h, 0.0f, 0.0f);
            }
            WalkRec w;
            w.px = (double)p.ix + (double)p.fx;
            w.py = (double)p.iy + (double)p.fy;
            if (kind == MI_KERNEL_TABLE) {
                // its weights are laid out over its cube-clipped footprint
                const int64_t stride = fj_hi - fj_lo + 1;
                w.px = __longlong_as_double((int64_t)table_offset(p) - (int64_t)fi_lo * stride - fj_lo);
                w.py = __longlong_as_double(stride);
            }
            w.fa = f.a;
            w.ij = ilo | (jlo << 16);
            w.kw = kind | (width << 8);
            cs.walk[lane] = w;
            if (kind == MI_KERNEL_GAUSS) {
                cs.gr[lane] = GaussRec{make_float4(f.b, f.t, f.n, 0.0f), f.e, 0.0};
            }
        }
    }
    if (lane == 0) {
        cs.off[0] = 0;
        cs.toff[0] = 0;
        cs.per = per;
        cs.rowmask = masks & 0xffffu;
        cs.chanmask = masks >> 16;
        cs.gmask = gauss;
    }
}

// The weight of a tabulated kernel at q^2 = dr^2 / h^2, as weight_at computes
// it, without branches: fb = 1/h^2, tab the kernel's table with its zero
// entry at LTAB_N - 1, CUBIC whether the seam term of the cubic spline
// applies.  The table position is clamped to that zero entry, so q >= 1 reads
// it (and the seam term is zero beyond q = 1/2): no mask, and the row is read
// for every pixel (no branch around the LDS load).
template <bool CUBIC>
__device__ inline float table_weight_select(const float4* __restrict__ tab, float q2, float fb) {
    const float q = __builtin_amdgcn_sqrtf(q2);
    const float x = fminf(q * (float)WTAB_N, (float)WTAB_N);
    const int i = (int)x;
    const float t = __builtin_amdgcn_fractf(x);
    const float4 c = tab[i];
    float w = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
    if (CUBIC) {  // R2 sqrt(1 - R2) / (8 * 1.5969), R2 = 4 q^2, zero for R2 >= 1
        w = fmaf(q2 * (float)(4.0 / (8.0 * 1.59689476201133)),
                 __builtin_amdgcn_sqrtf(fmaxf(fmaf(-4.0f, q2, 1.0f), 0.0f)), w);
    }
    return w * fb;
}

// Fills gt for the chunk's Gaussian particles (cs.gmask): thread t takes
// particle t % KC and rows (then columns) t / KC, t / KC + NT / KC, ... of its
// clipped footprint.  Loops end once no lane of the wave has a row (column)
// left.  The values are weight_at's: the same offsets and spans.
__device__ inline void fill_gauss_tables(const ChunkState& cs, GaussTables& gt) {
    const int k = threadIdx.x & (KC - 1), slot = threadIdx.x / KC;
    const bool gk = (cs.gmask >> k) & 1u;
    const WalkRec w = cs.walk[k];
    const float bf = cs.gr[k].f.x;
    const int i0 = w.ij & 0xffff, j0 = w.ij >> 16;
    const int rows = gk ? cs.ihi[k] - i0 + 1 : 0, cols = gk ? w.kw >> 8 : 0;
    for (int r = slot; __any(r < rows); r += NT / KC)
        if (r < rows) gt.ex[k][r] = erf_span((float)(w.px - (double)(i0 + r)) * bf, bf);
    for (int c = slot; __any(c < cols); c += NT / KC)
        if (c < cols) gt.ey[k][c] = erf_span((float)(w.py - (double)(j0 + c)) * bf, bf);
}

// The Gaussian kernel's weight as weight_at computes it, its separable
// factors ex, ey read from the tables: zero beyond the truncation edge, and
// the truncation term zmax = sqrt(t^2 - u^2) formed in FP64.
__device__ inline float gauss_weight(double dx, double dy, double fa, double fe, float t,
                                     float n, float ex, float ey) {
    const double dr2 = dx * dx + dy * dy;
    if (dr2 > fe) return 0.0f;
    const double z2 = (double)t * (double)t - dr2 * fa;
    const float zmax = z2 > 0.0 ? __builtin_amdgcn_sqrtf((float)z2) : 0.0f;
    return 0.25f * erff(zmax * 0.70710678118654752f) * ex * ey * n;
}

// Evaluates W for every (particle, pixel) pair of the chunk and hands each
// value to store(k, pixel-in-tile, w), or two of them to
// store2(k, pixel1, w1, pixel2, w2).  Particle k's pairs (pixel row l / width,
// column l % width of its clipped footprint, l < cnt) are split over threads
// toff[k] .. toff[k+1] - 1, `per` consecutive pairs each, so a thread stays on
// one particle: one search of the thread scan, then a walk over its pixels.  When every
// particle of the wave's pairs uses the item's LDS table, the walk is
// branch-free and takes two pixels per step, whose evaluations overlap.
template <class Store, class Store2>
__device__ inline void fill_weights(const ChunkState& cs, int x0, int y0,
                                    const float4* __restrict__ wtab, int tab_kind,
                                    const float4* __restrict__ tab, const GaussTables& gt,
                                    const float* __restrict__ uw, Store store, Store2 store2) {
    const int t = threadIdx.x;
    const bool active = t < cs.toff[KC];
    // k = the last particle with toff[k] <= t (toff is non-decreasing, and
    // particles without pairs share their successor's offset), in two rounds of
    // LDS reads: the eighth it lies in, then the entries of that eighth.
    static_assert(KC == 32, "the two-round search covers 32 particles");
    int k = 0;
    if (active) {
        const int m = 8 * ((cs.toff[8] <= t) + (cs.toff[16] <= t) + (cs.toff[24] <= t));
        const int4 u = *reinterpret_cast<const int4*>(cs.toff + m);
        const int4 v = *reinterpret_cast<const int4*>(cs.toff + m + 4);
        k = m + (u.y <= t) + (u.z <= t) + (u.w <= t) + (v.x <= t) + (v.y <= t) + (v.z <= t) +
            (v.w <= t);
    }
    const WalkRec w = cs.walk[k];
    const int width = w.kw >> 8, kind = w.kw & 255;
    const int l0 = (t - cs.toff[k]) * cs.per;
    const int l1 = active ? min(l0 + cs.per, cs.off[k + 1] - cs.off[k]) : l0;
    // l0 / width: (l0 + 1/2) / width lies at least 1/32 from an integer (l0 <
    // 256 pixels, width <= 16), far beyond the reciprocal's error
    int a = (int)(((float)l0 + 0.5f) * __builtin_amdgcn_rcpf((float)width));
    int b = l0 - a * width;
    const int ti = (w.ij & 0xffff) - x0, tj = (w.ij >> 16) - y0;  // footprint corner in the tile
    const double px = w.px - (double)x0, py = w.py - (double)y0;   // particle in tile coordinates

    if (tab_kind >= 0 && __all(!active || kind == tab_kind)) {
        // The particle's offset from the footprint's first pixel in FP32 (one
        // rounding of the exact FP64 value); q^2 follows in FP32, a few ulps
        // from the FP64 form, far below the table's own error.
        const float gx = (float)(px - (double)ti), gy = (float)(py - (double)tj);
        const float fa = (float)w.fa;
        // Pixel l of the footprint lies in row floor((l + 1/2) / width) (exact, as
        // for a above) and column l - row * width, all small integers formed in
        // FP32: no wrap tests, and the offsets from the particle need no
        // conversion.  Its index in the tile is row * TILE_Y + column + corner.
        const float rw = __builtin_amdgcn_rcpf((float)width), wf = (float)width;
        const float hrw = 0.5f * rw, corner4 = (float)(4 * (ti * TILE_Y + tj));
        auto at = [&](float lf, float& af, float& bf) {
            af = floorf(fmaf(lf, rw, hrw));  // (l + 1/2) / width, one rounding
            bf = fmaf(-af, wf, lf);
        };
        auto px_of = [&](float af, float bf) {
            const int q = (int)fmaf(af, (float)(4 * TILE_Y), fmaf(bf, 4.0f, corner4));  // 4 * index
            __builtin_assume((q & 3) == 0);
            return q >> 2;
        };
        auto walk = [&](auto cubic) {
            auto eval = [&](float af, float bf) {
                const float dx = gx - af, dy = gy - bf;
                return table_weight_select<decltype(cubic)::value>(tab, fmaf(dx, dx, dy * dy) * fa,
                                                                   fa);
            };
            float lf = (float)l0;
            const float lend = (float)l1;
            for (; lf + 1.0f < lend; lf += 2.0f) {
                float a1, b1, a2, b2;
                at(lf, a1, b1);
                at(lf + 1.0f, a2, b2);
                const float v1 = eval(a1, b1), v2 = eval(a2, b2);
                store2(k, px_of(a1, b1), v1, px_of(a2, b2), v2);
            }
            if (lf < lend) {
                float a1, b1;
                at(lf, a1, b1);
                store(k, px_of(a1, b1), eval(a1, b1));
            }
        };
        if (tab_kind == MI_KERNEL_CUBIC)  // uniform
            walk(std::true_type{});
        else
            walk(std::false_type{});
        return;
    }

    if (kind == MI_KERNEL_TABLE) {
        // the caller's weights (binning checked that they lie inside the table)
        const TableWalk tw = table_walk(w);
        for (int l = l0; l < l1; ++l) {
            const int i = x0 + ti + a, j = y0 + tj + b;
            store(k, (ti + a) * TILE_Y + tj + b, uw[tw.base + (int64_t)i * tw.stride + j]);
            if (++b == width) {
                b = 0;
                ++a;
            }
        }
        return;
    }
    if (kind == MI_KERNEL_GAUSS || !active) {
        // Gaussian particles: the separable factors come from the chunk's tables
        const double fe = cs.gr[k].e;
        const float4 gg = cs.gr[k].f;
        for (int l = l0; l < l1; ++l) {
            const int i = x0 + ti + a, j = y0 + tj + b;
            store(k, (ti + a) * TILE_Y + tj + b,
                  gauss_weight(w.px - (double)i, w.py - (double)j, w.fa, fe, gg.y, gg.z,
                               gt.ex[k][a], gt.ey[k][b]));
            if (++b == width) {
                b = 0;
                ++a;
            }
        }
        return;
    }
    const WFactors f{w.fa, 0.0, (float)w.fa, 0.0f, 0.0f};
    for (int l = l0; l < l1; ++l) {
        const int i = x0 + ti + a, j = y0 + tj + b;
        store(k, (ti + a) * TILE_Y + tj + b,
              weight_at(kind, w.px - (double)i, w.py - (double)j, f, wtab, tab_kind, tab));
        if (++b == width) {
            b = 0;
            ++a;
        }
    }
}

// Evaluates S for the chunk's KC particles x CB channels of the cell and
// hands the values to store2(k, channel1, s1, channel2, s2), two channels of
// one particle at a time; particles k >= nk and channels beyond the cube give
// 0.  Thread t takes particle t % KC and the channels t / KC + 8 i, so its
// record fields are read once.  Only the 16-channel blocks set in `blocks`
// (wave-uniform) are evaluated; the others are left untouched.  The values
// are line_flux's (for narrow spans the same series with its polynomial
// expanded, to FP32 rounding), computed as straight-line code (the window
// test is one unsigned compare and a select, the wide-channel erfc path one
// branch per block) so that the evaluations of a thread overlap.
template <int CB, class Store2>
__device__ inline void fill_spectra(const ChunkState& cs, int nk, int c0, int nc,
                                    const mi_cube_desc& d, Store2 store2,
                                    uint32_t blocks = (1u << (CB / 16)) - 1u) {
    constexpr int PER_BLOCK = 16 / (NT / KC);  // a thread's channels in one 16-channel block
    static_assert(PER_BLOCK == 2, "a thread stores its two channels of a block together");
    const int k = threadIdx.x & (KC - 1);
    const int cl0 = threadIdx.x / KC;
    const mi_particle& p = cs.p[k];
    const bool live = k < nk;
    const int lo = live ? p.c_lo : 1, hi = live ? min(p.c_hi, nc - 1) : 0;
    const int k0 = p.k0;
    const float fu = p.fu, s = p.s, amp = p.amp;
    const int spectrum = d.spectrum;
    // Channel cl0 + o of the cell lies in the window [lo, hi] when
    // (unsigned)(o - w0) <= wl; an empty window takes an offset no o reaches.
    const int w0 = lo <= hi ? lo - c0 - cl0 : (1 << 30);
    const uint32_t wl = lo <= hi ? (uint32_t)(hi - lo) : 0u;
    auto inside = [&](int o) { return (uint32_t)(o - w0) <= wl; };
    auto put = [&](int b, float v0, float v1) {
        constexpr int STEP = NT / KC;
        const int o0 = STEP * (b * PER_BLOCK), o1 = o0 + STEP;
        store2(k, cl0 + o0, inside(o0) ? v0 : 0.0f, cl0 + o1, inside(o1) ? v1 : 0.0f);
    };
    if (spectrum == MI_SPECTRUM_TABLE) {  // uniform: the caller's spectra
#pragma unroll
        for (int b = 0; b < CB / 16; ++b) {
            if (!((blocks >> b) & 1u)) continue;
            // the row of the caller's table from the thread's first channel on
            // (channel c0 + cl0 + o is inside the window when inside(o))
            const float* row = d.spectra + ((int64_t)k0 * d.spectra_ld + c0 + cl0);
            const int o0 = (NT / KC) * (b * PER_BLOCK), o1 = o0 + NT / KC;
            store2(k, cl0 + o0, inside(o0) ? amp * row[o0] : 0.0f, cl0 + o1,
                   inside(o1) ? amp * row[o1] : 0.0f);
        }
        return;
    }
    if (spectrum == MI_SPECTRUM_DIRAC) {  // uniform: the whole amplitude in the window
#pragma unroll
        for (int b = 0; b < CB / 16; ++b)
            if ((blocks >> b) & 1u) put(b, amp, amp);
        return;
    }
    // Narrow spans: span_series with its D-dependent factors folded once per
    // particle (D = s), 0.5 amp (erf(m + D/2) - erf(m - D/2)) =
    // exp(-m^2) (e0 + e1 m^2 + e2 m^4 + e3 m^6); the same polynomial expanded.
    const float D2 = s * s, K = amp * 0.564189583547756287f * s;  // amp D / sqrt(pi)
    const float e0 = K * (1.0f + D2 * (-1.0f / 12.0f + D2 * (1.0f / 160.0f - D2 * (1.0f / 2688.0f))));
    const float e1 = K * D2 * (1.0f / 6.0f + D2 * (-1.0f / 40.0f + D2 * (1.0f / 448.0f)));
    const float e2 = K * D2 * D2 * (1.0f / 120.0f - D2 * (1.0f / 672.0f));
    const float e3 = K * D2 * D2 * D2 * (1.0f / 5040.0f);
    // Channel c = c0 + cl0 + 8 j of this thread sits at k0 - c + fu - 1/2 =
    // A - 8 j from the line centre (in channels), A formed once; subtracting
    // the small integer 8 j from A is exact in the in-window range.
    const int kc = k0 - c0 - cl0;
    const float A = (float)kc + (fu - 0.5f);
    // exp(-m^2) as exp2(-n^2) with n = m sqrt(log2 e): n = sn (A - 8 j), and the
    // polynomial in m^2 = n^2 / log2 e with its coefficients scaled once.
    constexpr float LOG2E = 1.44269504088896341f;
    const float sn = s * 1.20112240878645110f;  // sqrt(log2 e)
    const float g1 = e1 * (1.0f / LOG2E), g2 = e2 * (1.0f / (LOG2E * LOG2E)),
                g3 = e3 * (1.0f / (LOG2E * LOG2E * LOG2E));
#pragma unroll
    for (int b = 0; b < CB / 16; ++b) {
        if (!((blocks >> b) & 1u)) continue;  // uniform
        float v[PER_BLOCK];
#pragma unroll
        for (int i = 0; i < PER_BLOCK; ++i) {
            const float n = sn * (A - (float)((NT / KC) * (b * PER_BLOCK + i))), n2 = n * n;
            v[i] = __builtin_amdgcn_exp2f(-n2) * fmaf(fmaf(fmaf(g3, n2, g2), n2, g1), n2, e0);
        }
        if (s > 0.5f) {
            // wide channels: a line this narrow spans a few channels, and the
            // erfc differences are taken only inside its window
#pragma unroll
            for (int i = 0; i < PER_BLOCK; ++i) {
                const int o = (NT / KC) * (b * PER_BLOCK + i);
                if (!inside(o)) continue;
                const float m = s * (((float)(kc - o) + fu) - 0.5f);  // kc - o = k0 - channel
                v[i] = 0.5f * amp * erf_diff(m - 0.5f * s, m + 0.5f * s);
            }
        }
        put(b, v[0], v[1]);
    }
}

// Copies the table of the kernel kind of the item's first particle (when it
// is a tabulated kind) into LDS, with a zero entry after it; returns that
// kind, or -1.  Particles of one
// cell are sorted with the tabulated kinds first, so this is the table the
// item uses (a second tabulated kind in the same cell reads the global copy).
// The copy is complete after the next __syncthreads().
__device__ inline int stage_table(const mi_particle* __restrict__ P,
                                  const uint32_t* __restrict__ order, int begin,
                                  const float4* __restrict__ wtab, float4* local) {
    const int kind = kind_of(P[order[begin]].r_kind);
    if (kind > MI_KERNEL_QUARTIC) return -1;
    for (int e = threadIdx.x; e < WTAB_N; e += NT) local[e] = wtab[kind * WTAB_N + e];
    if (threadIdx.x == 0) local[WTAB_N] = make_float4(0.f, 0.f, 0.f, 0.f);
    return kind;
}

// Cell coordinates of a work item.
struct CellPos {
    int x0, y0, c0;
};

__device__ inline CellPos cell_pos(int cell, const Grid& g) {
    const int cb = cell % g.ncb;
    const int txy = cell / g.ncb;
    return CellPos{(txy / g.nty) << g.tx_shift, (txy % g.nty) << g.ty_shift, cb << g.cb_shift};
}

}  // namespace mi
