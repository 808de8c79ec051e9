// Beam convolution of the resident cube: every channel map convolved with the
// beam image (martini/martini.py:863-901, which runs
// scipy.signal.fftconvolve(spatial_slice, beam.kernel, mode="same") per
// channel).  'same' mode with zero boundary, for an odd kx x ky beam image K:
//
//   out[i, j, c] = sum_{a < kx, b < ky} K[a, b] * in[i + kx/2 - a, j + ky/2 - b, c]
//
// The cube is [nx][ny][nc], channels contiguous, so a thread carries a vector
// of 4 channels through all taps.  A work group computes a 16 x 16 pixel x
// 4-channel output tile; the beam image is taken in blocks of H rows x W
// columns whose input halo ((16 + H - 1) x (16 + W - 1) pixels x 4 channels)
// and taps (H x W) fit in LDS, so any beam size runs with the same code.
// Direct summation in a fixed order (deterministic): FP32 along each row
// segment of a block, FP64 across them.  A separable beam (a circular or
// axis-aligned Gaussian's image is the outer product of two vectors) is
// convolved by the caller as two such passes, a kx x 1 and a 1 x ky image;
// those one-dimensional images take k_convolve_line, which reads each pixel's
// channels contiguously.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/martini_mi355x.h"
#include "mi_internal.h"

namespace mi {

constexpr int BT = 16;               // output tile edge (pixels)
constexpr int BEAM_LDS = 56 * 1024;  // bytes of LDS for one block's halo and taps
constexpr int BEAM_MAX_W = 128;      // beam columns per block

// Columns and rows of the beam image per block.
__host__ __device__ inline int beam_block_cols(int ky) { return ky < BEAM_MAX_W ? ky : BEAM_MAX_W; }

__host__ __device__ inline int beam_block_rows(int kx, int ky) {
    const int w = beam_block_cols(ky);
    int h = BEAM_LDS / (16 * (BT - 1 + w) + 4 * w) - (BT - 1);
    if (h > kx) h = kx;
    return h < 1 ? 1 : h;
}

__global__ __launch_bounds__(256) void k_convolve_beam(const float* __restrict__ in,
                                                       float* __restrict__ out, int nx, int ny,
                                                       int nc, const float* __restrict__ K,
                                                       int kx, int ky, int band_h, int band_w) {
    extern __shared__ float4 lds[];
    const int tiles_y = (ny + BT - 1) / BT;
    const int tile = blockIdx.x;
    const int x0 = (tile / tiles_y) * BT, y0 = (tile % tiles_y) * BT;
    const int c0 = blockIdx.y * 4;
    const int t = threadIdx.x;
    const int tx = t / BT, ty = t % BT;
    const int hx = kx / 2, hy = ky / 2;
    const int wy = BT - 1 + band_w;  // halo width along y
    float4* halo = lds;
    float* taps = reinterpret_cast<float*>(lds + (BT - 1 + band_h) * wy);

    // each row segment of a block is summed in FP32 and added to FP64 totals,
    // so the FP32 rounding grows with the block width, not with kx * ky
    double ax = 0.0, ay = 0.0, az = 0.0, aw = 0.0;
    for (int a0 = 0; a0 < kx; a0 += band_h) {
        const int H = min(band_h, kx - a0);
        for (int b0 = 0; b0 < ky; b0 += band_w) {
            const int W = min(band_w, ky - b0);
            // output (x0 + tx, y0 + ty) reads input (x0 + tx + hx - a, y0 + ty + hy - b)
            // for a in [a0, a0 + H), b in [b0, b0 + W)
            const int rx0 = x0 + hx - a0 - (H - 1);  // input row of halo row 0
            const int ry0 = y0 + hy - b0 - (W - 1);  // input column of halo column 0
            const int hrows = BT - 1 + H, hcols = BT - 1 + W;
            __syncthreads();  // the previous block's halo is no longer read
            for (int e = t; e < hrows * hcols; e += 256) {
                const int r = e / hcols, q = e - r * hcols;
                const int i = rx0 + r, j = ry0 + q;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (i >= 0 && i < nx && j >= 0 && j < ny) {
                    const float* src = in + ((size_t)i * ny + j) * nc + c0;
                    if (c0 + 3 < nc && (nc & 3) == 0) {
                        v = *reinterpret_cast<const float4*>(src);
                    } else {
                        v.x = src[0];
                        if (c0 + 1 < nc) v.y = src[1];
                        if (c0 + 2 < nc) v.z = src[2];
                        if (c0 + 3 < nc) v.w = src[3];
                    }
                }
                halo[r * wy + q] = v;
            }
            for (int e = t; e < H * W; e += 256) {
                const int da = e / W, db = e - da * W;
                taps[e] = K[(size_t)(a0 + da) * ky + b0 + db];
            }
            __syncthreads();
            // tap (a0 + da, b0 + db) reads halo row tx + H - 1 - da, column ty + W - 1 - db
            for (int da = 0; da < H; ++da) {
                const float4* hrow = halo + (tx + H - 1 - da) * wy + ty + W - 1;
                const float* krow = taps + da * W;
                float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
                for (int db = 0; db < W; ++db) {
                    const float k = krow[db];
                    const float4 v = hrow[-db];
                    acc.x = fmaf(k, v.x, acc.x);
                    acc.y = fmaf(k, v.y, acc.y);
                    acc.z = fmaf(k, v.z, acc.z);
                    acc.w = fmaf(k, v.w, acc.w);
                }
                ax += acc.x;
                ay += acc.y;
                az += acc.z;
                aw += acc.w;
            }
        }
    }
    const int i = x0 + tx, j = y0 + ty;
    if (i < nx && j < ny) {
        float* dst = out + ((size_t)i * ny + j) * nc + c0;
        dst[0] = (float)ax;
        if (c0 + 1 < nc) dst[1] = (float)ay;
        if (c0 + 2 < nc) dst[2] = (float)az;
        if (c0 + 3 < nc) dst[3] = (float)aw;
    }
}

// One-dimensional pass over a [O][L][E] view of the cube (E contiguous, in
// units of V: float4 or float), along L:
//
//   out[o][l][e] = sum_{a < k} K[a] * in[o][l + k/2 - a][e],  zero outside [0, L).
//
// A kx x 1 image is the pass along x (O = 1, L = nx, E = ny * nc), a 1 x ky
// image the pass along y (O = nx, L = ny, E = nc).  The 64 lanes of a wave
// take 64 consecutive elements of E (one 1 KB row of float4s per load), and
// each wave LINE_R consecutive outputs along L: it walks the LINE_R + k - 1
// input rows once, adding each to the outputs whose taps reach it, so an
// input row is loaded (LINE_R + k - 1) / LINE_R times per wave instead of k.
// The taps are held in LDS as FP64 and every sum is formed in FP64 in a fixed
// order (deterministic; FP32 rounding only at the output).
constexpr int LINE_R = 8;        // outputs per wave along the axis
constexpr int LINE_WAVES = 4;    // waves per work group (consecutive strips along L)
constexpr int LINE_MAX_K = 4096; // taps held in LDS (32 KB); longer images take k_convolve_beam
constexpr int LINE_G = 4;        // input rows loaded together

template <class V>
__device__ inline void fma_into(double (&acc)[4], double w, const V& v);
template <>
__device__ inline void fma_into<float4>(double (&acc)[4], double w, const float4& v) {
    acc[0] = fma(w, (double)v.x, acc[0]);
    acc[1] = fma(w, (double)v.y, acc[1]);
    acc[2] = fma(w, (double)v.z, acc[2]);
    acc[3] = fma(w, (double)v.w, acc[3]);
}
template <>
__device__ inline void fma_into<float>(double (&acc)[4], double w, const float& v) {
    acc[0] = fma(w, (double)v, acc[0]);
}
__device__ inline float4 round_out(const double (&a)[4], float4*) {
    return make_float4((float)a[0], (float)a[1], (float)a[2], (float)a[3]);
}
__device__ inline float round_out(const double (&a)[4], float*) { return (float)a[0]; }

template <class V>
__global__ __launch_bounds__(64 * LINE_WAVES) void k_convolve_line(
    const V* __restrict__ in, V* __restrict__ out, int L, int64_t E, const float* __restrict__ K,
    int k) {
    extern __shared__ double taps[];
    for (int a = threadIdx.x; a < k; a += 64 * LINE_WAVES) taps[a] = (double)K[a];
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l0 = (blockIdx.y * LINE_WAVES + wave) * LINE_R;
    if (e >= E || l0 >= L) return;
    const V* src = in + (int64_t)blockIdx.z * L * E + e;
    V* dst = out + (int64_t)blockIdx.z * L * E + e;
    double acc[LINE_R][4];
#pragma unroll
    for (int q = 0; q < LINE_R; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] = 0.0;
    // Input row m0 + r feeds output l0 + q through tap a = q + k - 1 - r; rows
    // outside [0, L) are zero and skipped.  Rows are taken LINE_G at a time, their
    // loads issued together.
    const int m0 = l0 + k / 2 - (k - 1);
    const int r0 = max(0, -m0), r1 = min(LINE_R + k - 1, L - m0);
    int r = r0;
    for (; r + LINE_G <= r1; r += LINE_G) {
        V v[LINE_G];
#pragma unroll
        for (int g = 0; g < LINE_G; ++g) v[g] = src[(int64_t)(m0 + r + g) * E];
#pragma unroll
        for (int g = 0; g < LINE_G; ++g)
#pragma unroll
            for (int q = 0; q < LINE_R; ++q) {
                const int a = q + k - 1 - (r + g);
                if (a >= 0 && a < k) fma_into<V>(acc[q], taps[a], v[g]);  // uniform
            }
    }
    for (; r < r1; ++r) {
        const V v = src[(int64_t)(m0 + r) * E];
#pragma unroll
        for (int q = 0; q < LINE_R; ++q) {
            const int a = q + k - 1 - r;
            if (a >= 0 && a < k) fma_into<V>(acc[q], taps[a], v);  // uniform
        }
    }
#pragma unroll
    for (int q = 0; q < LINE_R; ++q)
        if (l0 + q < L) dst[(int64_t)(l0 + q) * E] = round_out(acc[q], (V*)nullptr);
}

static hipError_t convolve_line(const float* in, float* out, int64_t O, int L, int64_t E_floats,
                                const float* K, int k, hipStream_t s) {
    const unsigned strips = (unsigned)((L + LINE_R * LINE_WAVES - 1) / (LINE_R * LINE_WAVES));
    const size_t lds = sizeof(double) * (size_t)k;
    const bool vec = E_floats % 4 == 0 &&
                     ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    if (vec) {
        const int64_t E = E_floats / 4;
        hipLaunchKernelGGL(k_convolve_line<float4>, dim3((unsigned)((E + 63) / 64), strips, (unsigned)O),
                           dim3(64 * LINE_WAVES), lds, s, reinterpret_cast<const float4*>(in),
                           reinterpret_cast<float4*>(out), L, E, K, k);
    } else {
        hipLaunchKernelGGL(k_convolve_line<float>, dim3((unsigned)((E_floats + 63) / 64), strips, (unsigned)O),
                           dim3(64 * LINE_WAVES), lds, s, in, out, L, E_floats, K, k);
    }
    return hipGetLastError();
}

size_t beam_lds_bytes(int kx, int ky) {
    const int h = beam_block_rows(kx, ky), w = beam_block_cols(ky);
    return sizeof(float4) * (size_t)(BT - 1 + h) * (BT - 1 + w) + sizeof(float) * (size_t)h * w;
}

hipError_t convolve_beam(const float* in, float* out, int nx, int ny, int nc, const float* K,
                         int kx, int ky, hipStream_t s) {
    if (ky == 1 && kx <= LINE_MAX_K)  // along x: [1][nx][ny * nc]
        return convolve_line(in, out, 1, nx, (int64_t)ny * nc, K, kx, s);
    if (kx == 1 && ky <= LINE_MAX_K)  // along y: [nx][ny][nc]
        return convolve_line(in, out, nx, ny, nc, K, ky, s);
    const int h = beam_block_rows(kx, ky), w = beam_block_cols(ky);
    const size_t lds = beam_lds_bytes(kx, ky);
    const unsigned tiles = (unsigned)(((nx + BT - 1) / BT) * ((ny + BT - 1) / BT));
    const unsigned cq = (unsigned)((nc + 3) / 4);
    hipLaunchKernelGGL(k_convolve_beam, dim3(tiles, cq), dim3(256), lds, s, in, out, nx, ny, nc, K,
                       kx, ky, h, w);
    return hipGetLastError();
}

}  // namespace mi
