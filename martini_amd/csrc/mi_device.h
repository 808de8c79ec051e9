// Device math shared by the insertion kernels: SPH kernel integrals over a
// pixel and the per-channel line profile.
//
// Kernel integrals restate martini/sph_kernels.py (cubic :791-858,
// WendlandC2 :400-441, WendlandC6 :544-685, quartic :1486-1564, Gaussian
// :994-1044, Dirac :1136-1165).  The closed forms sum terms ~30x larger than
// their result, so they are evaluated once in FP64 into an interpolation table
// over q = dr/h, which the insertion kernels read.  The Gaussian kernel and the
// spectra use FP32 differences of erf over one pixel or channel, taken as a
// midpoint series for narrow spans and on the small-erfc side otherwise, so
// neither cancels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/martini_mi355x.h"

namespace mi {

// One cell per (pixel tile, channel block).  A tile row (16 pixels along y)
// is one operand row of the products.  Two shapes, chosen per cube by its
// channel count (wide_cells): 8 x 16-pixel tiles x 128-channel blocks, so a
// particle's W is formed once per 128 channels of its line, for cubes of 128
// channels or more; 16 x 16-pixel tiles x 64-channel blocks below that, where
// a 128-channel block would be half empty.  Both hold 16384 voxels.
template <int TX_, int TY_, int CB_>
struct CellGeo {
    static constexpr int TX = TX_;  // tile rows (pixels along x)
    static constexpr int TY = TY_;  // tile columns (pixels along y)
    static constexpr int PX = TX_ * TY_;
    static constexpr int CB = CB_;  // channels per block
};
using GeoSquare = CellGeo<16, 16, 64>;
using GeoWide = CellGeo<8, 16, 128>;
constexpr int CELL_VOXELS = 16384;  // one cell's (partial) tile, either shape
static_assert(GeoSquare::PX * GeoSquare::CB == CELL_VOXELS &&
                  GeoWide::PX * GeoWide::CB == CELL_VOXELS,
              "both cell shapes hold the same number of voxels");
// A cell's sub-tiles: one tile row (16 pixels) x SUB_CH channels each, 64 in
// either shape.  A partial tile stores them one after another, each 1 KiB
// ([row][channel block][pixel][SUB_CH]).  Bit sub_bit(row, block) of its mask
// is set where the item wrote that sub-tile; the others hold nothing of this
// call and are read as zero.  The bits follow the MFMA gather's waves: wave
// w = row / 4 + (TX / 4) * (block / 4) owns bits 16 w .. 16 w + 15 (4 (row % 4)
// + block % 4), so each wave stores its own 16 bits.
constexpr int SUB_CH = 16;
static_assert(GeoSquare::TX * (GeoSquare::CB / SUB_CH) == 64 &&
                  GeoWide::TX * (GeoWide::CB / SUB_CH) == 64,
              "one 64-bit sub-tile mask per partial tile");
__host__ __device__ constexpr int sub_offset(int cb_blocks, int row, int block, int pixel) {
    return ((row * cb_blocks + block) * 16 + pixel) * 16;
}
__host__ __device__ constexpr int sub_bit(int tx_shift, int row, int block) {
    return 16 * ((row >> 2) + ((1 << tx_shift) >> 2) * (block >> 2)) + (row & 3) * 4 + (block & 3);
}
constexpr bool sub_bits_distinct(int tx_shift, int rows, int blocks) {
    uint64_t seen = 0u;
    for (int r = 0; r < rows; ++r)
        for (int b = 0; b < blocks; ++b) {
            const int bit = sub_bit(tx_shift, r, b);
            if (bit < 0 || bit > 63 || ((seen >> bit) & 1u)) return false;
            seen |= 1ull << bit;
        }
    return true;
}
static_assert(sub_bits_distinct(3, GeoWide::TX, GeoWide::CB / SUB_CH) &&
                  sub_bits_distinct(4, GeoSquare::TX, GeoSquare::CB / SUB_CH),
              "one mask bit per sub-tile");
constexpr int TILE_MAX = 16;  // the longer tile side of either shape
constexpr int TILE_Y = 16;    // tile columns, the same in both shapes
static_assert(GeoSquare::TY == TILE_Y && GeoWide::TY == TILE_Y, "one row length");
__host__ __device__ inline bool wide_cells(int nc) { return nc >= 128; }
// Particles per cell processed by one work item: longer cells are split and
// their partial tiles summed in item order by a reduction pass.  The split
// length depends only on the pair count (about ITEM_TARGET items, between
// ITEM_MIN and ITEM_MAX particles each), so a given input always sums in the
// same order.
constexpr int ITEM_MIN = 256;
constexpr int ITEM_MAX = 4096;
constexpr int ITEM_TARGET = 2048;
// Bits of the launch-order key of an item (ITEM_MAX - length, or the padding
// key 2^(ITEM_KEY_BITS - 1) of slots past the item count).
constexpr int ITEM_KEY_BITS = 15;
static_assert(ITEM_MAX < (1 << (ITEM_KEY_BITS - 1)), "item lengths fit the launch-order key");

__host__ __device__ inline int item_length(int64_t n_pairs) {
    int64_t len = (n_pairs / ITEM_TARGET + 31) / 32 * 32;
    return (int)(len < ITEM_MIN ? ITEM_MIN : (len > ITEM_MAX ? ITEM_MAX : len));
}
// Key layout: cell << KEY_LOCAL_BITS | kernel group << 8 | 8-bit local order
// (zero, keeping the record order, or a hash of the particle index).
constexpr int KEY_LOCAL_BITS = 10;

__host__ __device__ inline int kind_of(int32_t r_kind) { return r_kind & 15; }
// MI_KERNEL_TABLE particles carry the offset of their weights in the h field.
__host__ __device__ inline uint32_t table_offset(const mi_particle& p) {
    uint32_t u;
    __builtin_memcpy(&u, &p.h, 4);
    return u;
}
__host__ __device__ inline int gidx_of(int32_t r_kind) { return (r_kind >> 4) & 3; }
__host__ __device__ inline int range_of(int32_t r_kind) { return r_kind >> 8; }

// Clipped footprint of a particle: pixel rows [i_lo, i_hi] and columns
// [j_lo, j_hi] with |i - x| <= r (martini/martini.py:271-274).  Returns false
// when empty.  x = ix + fx with fx in [0, 1), r an integer, so
// ceil(x - r) = ix - r + (fx > 0) and floor(x + r) = ix + r exactly.
//
// The tabulated kernels (MI_KERNEL_CUBIC .. MI_KERNEL_QUARTIC) weigh a pixel
// zero when q = |pixel - particle| / h >= 1, so their footprint is further cut
// to the rows and columns within h (1 + 2^-12) of the particle: every pixel
// left out has q^2 > 1 + 2^-12, which the weight's FP32 evaluation of q^2
// (a few ulps) cannot bring below 1.  This drops none of the weights the
// square mask gives and about a fifth of the pixels on typical discs.  The
// bounds are offsets from ix (iy) formed in FP32 from the fraction fx (fy),
// whose rounding (below 2^-23 of h) the margin covers, without contraction,
// so host and device agree.
__host__ __device__ inline bool footprint(const mi_particle& p, int nx, int ny, int& i_lo,
                                          int& i_hi, int& j_lo, int& j_hi) {
#pragma clang fp contract(off)
    int r = range_of(p.r_kind);
    if (r < 0) return false;
    i_lo = p.ix - r + (p.fx > 0.0f ? 1 : 0);
    i_hi = p.ix + r;
    j_lo = p.iy - r + (p.fy > 0.0f ? 1 : 0);
    j_hi = p.iy + r;
    if (kind_of(p.r_kind) <= MI_KERNEL_QUARTIC && p.h > 0.0f && p.h <= (float)r) {
        const float R = p.h * (1.0f + 0x1p-12f);
        const int a = p.ix + (int)ceilf(p.fx - R), b = p.ix + (int)floorf(p.fx + R);
        const int c = p.iy + (int)ceilf(p.fy - R), d = p.iy + (int)floorf(p.fy + R);
        if (a > i_lo) i_lo = a;
        if (b < i_hi) i_hi = b;
        if (c > j_lo) j_lo = c;
        if (d < j_hi) j_hi = d;
    }
    if (i_lo < 0) i_lo = 0;
    if (j_lo < 0) j_lo = 0;
    if (i_hi > nx - 1) i_hi = nx - 1;
    if (j_hi > ny - 1) j_hi = ny - 1;
    return i_lo <= i_hi && j_lo <= j_hi;
}

// ---------------------------------------------------------------- kernels
__host__ __device__ inline double w_cubic(double dx, double dy, double h) {
    double R2 = 4.0 * (dx * dx + dy * dy) / (h * h), F;
    if (R2 == 0.0) {
        F = 0.75;
    } else if (R2 <= 1.0) {
        double A = sqrt(1.0 - R2), B = sqrt(4.0 - R2);
        double lA = log((1.0 + A) / sqrt(R2));
        double I1 = A - 0.5 * A * A * A - 1.5 * R2 * A + (3.0 / 32.0) * A * (3.0 * R2 + 2.0) +
                    (9.0 / 32.0) * R2 * R2 * lA;
        double I3 = -B * (3.0 * R2 + 56.0) * 0.25 + A * (4.0 * R2 + 50.0) * 0.125 -
                    0.375 * R2 * (R2 + 16.0) * log((2.0 + B) / (1.0 + A)) +
                    2.0 * (3.0 * R2 + 4.0) * (B - A) + 2.0 * (B * B * B - A * A * A);
        F = I1 + 0.25 * I3;
    } else if (R2 <= 4.0) {
        double B = sqrt(4.0 - R2);
        double I2 = -B * (3.0 * R2 + 56.0) * 0.25 -
                    0.375 * R2 * (R2 + 16.0) * log((2.0 + B) / sqrt(R2)) +
                    2.0 * (3.0 * R2 + 4.0) * B + 2.0 * B * B * B;
        F = 0.25 * I2;
    } else {
        return 0.0;
    }
    return F * (4.0 / 1.59689476201133) / (h * h);
}

__host__ __device__ inline double w_wc2(double dx, double dy, double h) {
    double R2 = (dx * dx + dy * dy) / (h * h), G;
    if (R2 == 0.0) {
        G = 2.0 / 3.0;
    } else if (R2 < 1.0) {
        double A = sqrt(1.0 - R2);
        G = 5.0 * R2 * R2 * (0.5 * R2 + 3.0) * log((1.0 + A) / sqrt(R2)) +
            A * (-13.5 * R2 * R2 - (14.0 / 3.0) * R2 + 2.0 / 3.0);
    } else {
        return 0.0;
    }
    return G * (21.0 / (2.0 * M_PI)) / (h * h);
}

// WendlandC6 indefinite integral in z at fixed R (the reference's rounded
// coefficients), grouped by the factor each term carries.
__host__ __device__ inline double wc6_indef(double R, double z) {
    double R2 = R * R, z2 = z * z;
    double S = sqrt(R2 + z2), L = log(S + z);
    double R4 = R2 * R2, R6 = R4 * R2, R8 = R4 * R4, R10 = R8 * R2, R12 = R6 * R6;
    double z3 = z2 * z, z5 = z3 * z2, z7 = z5 * z2, z9 = z7 * z2, z11 = z9 * z2;
    double plain = -231.0 * R10 * z - 385.0 * R8 * z3 - 1155.0 * R8 * z - 462.0 * R6 * z5 -
                   1540.0 * R6 * z3 - 462.0 * R6 * z - 330.0 * R4 * z7 - 1386.0 * R4 * z5 -
                   462.0 * R4 * z3 + 66.0 * R4 * z - (128.0 + 1.0 / 3.0) * R2 * z9 -
                   660.0 * R2 * z7 - 277.2 * R2 * z5 + 44.0 * R2 * z3 - 11.0 * R2 * z -
                   21.0 * z11 - (128.0 + 1.0 / 3.0) * z9 - 66.0 * z7 + 13.2 * z5 -
                   (11.0 / 3.0) * z3 + z;
    double withS = (8.0 / 3.0) * z11 + (16.0 + 4.0 / 15.0) * R2 * z9 + 70.4 * z9 +
                   360.8 * R2 * z7 + 132.0 * z7 + 550.0 * R2 * z5 + 24.7813 * R10 * z +
                   530.75 * R8 * z + 47.4792 * R8 * z3 + 767.25 * R6 * z + 58.0167 * R6 * z5 +
                   819.5 * R6 * z3 + 41.7 * R4 * z7 + 752.4 * R4 * z5 + 896.5 * R4 * z3;
    double withL = 7.21875 * R12 + 173.25 * R10 + 288.75 * R8;
    return plain + withS * S + withL * L;
}

__host__ __device__ inline double w_wc6(double dx, double dy, double h) {
    double R = sqrt(dx * dx + dy * dy) / h, v;
    const double norm = 1365.0 / (64.0 * M_PI);
    if (R == 0.0) v = norm * 2.0 * (4.0 / 15.0);
    else if (R < 1.0) v = norm * 2.0 * (wc6_indef(R, sqrt(1.0 - R * R)) - wc6_indef(R, 0.0));
    else return 0.0;
    return v / (h * h);
}

// Indefinite integral over z of (A - sqrt(R^2+z^2))^4 at fixed R.
__host__ __device__ inline double quartic_IA(double R, double z, double A) {
    double q = sqrt(z * z + R * R), R2 = R * R, A2 = A * A;
    return A2 * A2 * z - 2.0 * A2 * A * z * q + 2.0 * A2 * z * (3.0 * R2 + z * z) -
           0.5 * A * R2 * (4.0 * A2 + 3.0 * R2) * asinh(z / R) -
           0.5 * A * z * q * (5.0 * R2 + 2.0 * z * z) + R2 * R2 * z + (2.0 / 3.0) * R2 * z * z * z +
           0.2 * z * z * z * z * z;
}

__host__ __device__ inline double w_quartic(double dx, double dy, double h) {
    double R = sqrt(dx * dx + dy * dy) / h, v = 0.0;
    if (R == 0.0) {
        v = 384.0 / 3125.0;
    } else if (R < 1.0) {
        if (R < 0.2) v += 10.0 * quartic_IA(R, sqrt(0.04 - R * R), 0.2);
        if (R < 0.6) v -= 5.0 * quartic_IA(R, sqrt(0.36 - R * R), 0.6);
        v += quartic_IA(R, sqrt(1.0 - R * R), 1.0);
    } else {
        return 0.0;
    }
    return v * (2.0 * 15625.0 / (512.0 * M_PI)) / (h * h);
}

// erf(b) - erf(a) for a <= b, taken where the two terms do not cancel.
__device__ inline float erf_diff(float a, float b) {
    if (a >= 0.0f) return erfcf(a) - erfcf(b);
    if (b <= 0.0f) return erfcf(-b) - erfcf(-a);
    return erff(b) - erff(a);
}

// erf(m + D/2) - erf(m - D/2) for D >= 0.  Narrow spans use the midpoint
// Taylor series of the integral of exp(-t^2) (terms to D^7; error below 1e-7
// of the span's peak for D <= 0.5), which has no cancellation and costs one
// hardware exp (v_exp_f32 of -m^2 log2 e: relative error a few 1e-7 for the
// m^2 <= 25 that carry flux); wider spans take erf_diff.
__device__ inline float span_series(float m, float D) {
    const float m2 = m * m, D2 = D * D;
    float p = 64.0f * m2 - 480.0f;
    p = p * m2 + 720.0f;
    p = p * m2 - 120.0f;                                   // H6-type polynomial
    float q = (16.0f * m2 - 48.0f) * m2 + 12.0f;            // H4-type polynomial
    float poly = 1.0f + D2 * ((4.0f * m2 - 2.0f) * (1.0f / 24.0f) +
                              D2 * (q * (1.0f / 1920.0f) + D2 * p * (1.0f / 322560.0f)));
    return 1.12837916709551257f * __expf(-m2) * D * poly;
}

__device__ inline float erf_span(float m, float D) {
    if (D <= 0.5f) return span_series(m, D);
    return erf_diff(m - 0.5f * D, m + 0.5f * D);
}

constexpr double GAUSS_SIG = 0.42466090014400953;  // 1 / (2 sqrt(2 ln 2)): FWHM = 1

// The four closed-form kernels depend on q = dr/h only: w = F(q) / h^2 with
// F(q) = w(q, 0, 1).  F is tabulated at WTAB_N intervals of [0, 1] from the
// FP64 closed forms above; entry i holds the 4-point Lagrange cubic through
// the nodes i-1 .. i+2 (F(-q) = F(q)) in monomial form, evaluated by Horner.
// WTAB_N = 250 puts the kernels' seams (cubic q = 1/2, quartic q = 1/5 and
// 3/5) on nodes; the error, about dq^4 max|F^(4)| / 24, stays below 2e-8 of
// F(0), and one kernel's table (4 KB) fits in LDS beside the tiles.  Node WTAB_N holds the left limit F(1-): WendlandC6's rounded
// coefficients leave it nonzero (~4e-5 of F(0)), so its node beyond 1 is
// extrapolated from inside; the others vanish smoothly at q = 1.  q >= 1
// gives 0, as the reference's strict or zero-valued support edge does.
constexpr int WTAB_N = 250;
constexpr int WTAB_KERNELS = 4;  // indexed by kernel kind: cubic, WC2, WC6, quartic

// The reference's cubic I3 exceeds the exact outer-shell integral by R2*A/8
// (R2 = 4 q^2 <= 1, A = sqrt(1 - R2)); with its 1/4 and normalisation this adds
// R2*A/(8 * 1.59689476201133) to the weight.  Its square-root behaviour at
// the seam q = 1/2 is kept out of the table and added analytically.
__host__ __device__ inline double cubic_seam_term(double q) {
    const double R2 = 4.0 * q * q;
    return R2 <= 1.0 ? R2 * sqrt(1.0 - R2) / (8.0 * 1.59689476201133) : 0.0;
}

__host__ __device__ inline double wtab_eval(int which, double q) {
    switch (which) {
    case MI_KERNEL_CUBIC: return w_cubic(q, 0.0, 1.0) - cubic_seam_term(q);
    case MI_KERNEL_WC2: return w_wc2(q, 0.0, 1.0);
    case MI_KERNEL_WC6: return w_wc6(q, 0.0, 1.0);
    default: return w_quartic(q, 0.0, 1.0);
    }
}

__host__ __device__ inline double wtab_node(int which, int node) {
    const int a = node < 0 ? -node : node;
    if (a < WTAB_N) return wtab_eval(which, (double)a / WTAB_N);
    const double left = wtab_eval(which, 1.0 - 1e-13);  // F(1-)
    if (a == WTAB_N) return left;
    if (which != MI_KERNEL_WC6) return 0.0;
    return 3.0 * left - 3.0 * wtab_eval(which, (double)(WTAB_N - 1) / WTAB_N) +
           wtab_eval(which, (double)(WTAB_N - 2) / WTAB_N);
}

__device__ inline float table_weight(const float4* __restrict__ tab, float q) {
    if (!(q < 1.0f)) return 0.0f;
    const float x = q * (float)WTAB_N;
    const int i = min((int)x, WTAB_N - 1);
    const float t = x - (float)i;
    const float4 c = tab[i];  // the interval's cubic, c.x + c.y t + c.z t^2 + c.w t^3
    return fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
}
// Per-particle factors of the pixel weight, formed once per particle so the
// per-pixel evaluation has no division:
//   table kernels: a = 1/h^2 (FP64, for q^2), b = (float)(1/h^2);
//   Gaussian:      a = 1/(h sig)^2, e = (t h sig + 1/sqrt 2)^2 (the truncation
//                  edge of sph_kernels.py:1041, squared), b = 1/(h sig sqrt 2)
//                  (one pixel in erf-argument units), t and n = 1/norm_t.
struct WFactors {
    double a, e;
    float b, t, n;
};

// 1 / x in FP64: the hardware reciprocal refined by two Newton steps (within
// an ulp of the quotient for finite nonzero x); for x = 0 or infinite the
// hardware value (an infinity or a zero) is the quotient itself.
__device__ inline double recip(double x) {
    const double r0 = __builtin_amdgcn_rcp(x);
    double r = fma(r0, fma(-x, r0, 1.0), r0);
    r = fma(r, fma(-x, r, 1.0), r);
    return (x == 0.0 || __builtin_isinf(x)) ? r0 : r;
}

__device__ inline WFactors weight_factors(int kind, float h, float gt, float ginv) {
    WFactors f{0.0, 0.0, 0.0f, gt, ginv};
    if (kind <= MI_KERNEL_QUARTIC) {
        // 1/h^2 within two FP32 ulps (its uses round q^2 to FP32)
        f.b = __builtin_amdgcn_rcpf(h * h);
        f.a = (double)f.b;
    } else if (kind == MI_KERNEL_GAUSS) {
        const double hs = (double)h * GAUSS_SIG;
        const double rh = recip(hs);
        f.a = rh * rh;
        const double edge = (double)gt * hs + 0.70710678118654752;
        f.e = edge * edge;
        f.b = (float)(rh * 0.70710678118654752);  // 1 / (hs sqrt 2)
    }
    return f;
}

// Weight of a particle at pixel offset (dx, dy) = particle - pixel centre,
// exact in FP64 from (ix - i) + fx: the integral of its kernel over the pixel
// per unit pixel area (sph_kernels.py:85-119 with h = h_px * rescale).
// wtab: WTAB_KERNELS x WTAB_N entries filled by fill_wtab.
//
// Gaussian: the offset and the truncation term zmax = sqrt(t^2 - u^2),
// u = dr/(h sig), are formed in FP64: zmax has a square-root singularity at
// the truncation radius, where the t = 3 kernel is still 1% of its peak, so
// FP32 rounding of u would show there at the 1e-5 level.
// `local`, when given, is a copy of kernel `local_kind`'s table (in LDS), read
// instead of wtab for that kind.
__device__ inline float weight_at(int kind, double dx, double dy, const WFactors& f,
                                  const float4* __restrict__ wtab, int local_kind = -1,
                                  const float4* __restrict__ local = nullptr) {
    const double dr2 = dx * dx + dy * dy;
    if (kind <= MI_KERNEL_QUARTIC) {
        const float q2 = (float)(dr2 * f.a);
        const float q = __builtin_amdgcn_sqrtf(q2);  // hardware sqrt, within 1 ulp
        float w = kind == local_kind ? table_weight(local, q)
                                     : table_weight(wtab + kind * WTAB_N, q);
        if (kind == MI_KERNEL_CUBIC && q2 <= 0.25f) {
            const float R2 = 4.0f * q2;
            w += R2 * __builtin_amdgcn_sqrtf(1.0f - R2) * (float)(1.0 / (8.0 * 1.59689476201133));
        }
        return w * f.b;
    }
    if (kind == MI_KERNEL_GAUSS) {
        if (dr2 > f.e) return 0.0f;
        const double z2 = (double)f.t * (double)f.t - dr2 * f.a;
        // z2 needs FP64 (t^2 - u^2 cancels); its root does not
        const float zmax = z2 > 0.0 ? __builtin_amdgcn_sqrtf((float)z2) : 0.0f;
        const float ex = erf_span((float)dx * f.b, f.b);
        const float ey = erf_span((float)dy * f.b, f.b);
        return 0.25f * erff(zmax * 0.70710678118654752f) * ex * ey * f.n;
    }
    if (kind == MI_KERNEL_DIRAC) return (fabs(dx) < 0.5 && fabs(dy) < 0.5) ? 1.0f : 0.0f;
    return 0.0f;
}

__device__ inline float pixel_weight(int kind, double dx, double dy, float h, float gt,
                                     float ginv, const float4* __restrict__ wtab) {
    return weight_at(kind, dx, dy, weight_factors(kind, h, gt, ginv), wtab);
}

// Fills the weight table: one thread per entry (WTAB_KERNELS * WTAB_N threads).
__device__ inline void fill_wtab(float4* __restrict__ tab) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= WTAB_KERNELS * WTAB_N) return;
    const int which = e / WTAB_N, i = e - which * WTAB_N;
    // Lagrange cubic through nodes i-1 .. i+2 (t = -1, 0, 1, 2) in monomial form
    const double fm = wtab_node(which, i - 1), f0 = wtab_node(which, i);
    const double f1 = wtab_node(which, i + 1), f2 = wtab_node(which, i + 2);
    tab[e] = make_float4((float)f0, (float)(-fm / 3.0 - f0 / 2.0 + f1 - f2 / 6.0),
                         (float)(fm / 2.0 - f0 + f1 / 2.0),
                         (float)(-fm / 6.0 + f0 / 2.0 - f1 / 2.0 + f2 / 6.0));
}
// Spectrum of a particle in channel c (Jy per unit weight): amp times the
// Gaussian line integrated over the channel (spectral_models.py:347-425), or
// amp inside the Dirac window (:510-570).  [lo, hi] is the particle's channel
// window, u = k0 + fu its line centre in channel units, s = |dv|/(sqrt 2 sigma).
__device__ inline float line_flux(int lo, int hi, int k0, float fu, float s, float amp, int c,
                                  int spectrum) {
    if (c < lo || c > hi) return 0.0f;
    if (spectrum == MI_SPECTRUM_DIRAC) return amp;
    const float dd = (float)(k0 - c) + fu;  // u - c: the channel spans s*(dd-1) .. s*dd
    return 0.5f * amp * erf_span(s * (dd - 0.5f), s);
}

// Spectrum value from the caller's table (MI_SPECTRUM_TABLE): amp times entry
// c of table row `row`, inside the particle's window [lo, hi].
__device__ inline float table_flux(int lo, int hi, int row, float amp, int c,
                                   const float* __restrict__ tab, int64_t ld) {
    if (c < lo || c > hi) return 0.0f;
    return amp * tab[(int64_t)row * ld + c];
}

__device__ inline float channel_flux(const mi_particle& p, int c, const mi_cube_desc& d) {
    if (d.spectrum == MI_SPECTRUM_TABLE)
        return table_flux(p.c_lo, p.c_hi, p.k0, p.amp, c, d.spectra, d.spectra_ld);
    return line_flux(p.c_lo, p.c_hi, p.k0, p.fu, p.s, p.amp, c, d.spectrum);
}

}  // namespace mi
