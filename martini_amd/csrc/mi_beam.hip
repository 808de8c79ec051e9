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
// convolved by the caller as two such passes, a kx x 1 and a 1 x ky image.
#include <hip/hip_runtime.h>

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

size_t beam_lds_bytes(int kx, int ky) {
    const int h = beam_block_rows(kx, ky), w = beam_block_cols(ky);
    return sizeof(float4) * (size_t)(BT - 1 + h) * (BT - 1 + w) + sizeof(float) * (size_t)h * w;
}

hipError_t convolve_beam(const float* in, float* out, int nx, int ny, int nc, const float* K,
                         int kx, int ky, hipStream_t s) {
    const int h = beam_block_rows(kx, ky), w = beam_block_cols(ky);
    const size_t lds = beam_lds_bytes(kx, ky);
    const unsigned tiles = (unsigned)(((nx + BT - 1) / BT) * ((ny + BT - 1) / BT));
    const unsigned cq = (unsigned)((nc + 3) / 4);
    hipLaunchKernelGGL(k_convolve_beam, dim3(tiles, cq), dim3(256), lds, s, in, out, nx, ny, nc, K,
                       kx, ky, h, w);
    return hipGetLastError();
}

}  // namespace mi
