// Beam convolution of the resident cube: every channel map convolved with the
// beam image (martini/martini.py:863-901, which runs
// scipy.signal.fftconvolve(spatial_slice, beam.kernel, mode="same") per
// channel).  'same' mode with zero boundary, for an odd kx x ky beam image K:
//
//   out[i, j, c] = sum_{a < kx, b < ky} K[a, b] * in[i + kx/2 - a, j + ky/2 - b, c]
//
// The cube is [nx][ny][nc], channels contiguous, so a thread carries a vector
// of 4 channels through all taps.  A work group computes a 16 x 16 pixel x
// 4-channel output tile; the beam's rows are taken in bands whose input halo
// ((16 + H - 1) x (16 + ky - 1) pixels x 4 channels) and taps (H x ky) fit in
// LDS, so any beam size runs with the same code.  Direct summation in a fixed
// order (deterministic): FP32 along each beam row, FP64 across rows.
#include <hip/hip_runtime.h>

#include "../../include/martini_mi355x.h"
#include "mi_internal.h"

namespace mi {

constexpr int BT = 16;            // output tile edge (pixels)
constexpr int BEAM_LDS = 56 * 1024;  // bytes of LDS for one band's halo and taps

// Rows of the beam image per band for a ky-wide beam.
__host__ __device__ inline int beam_band_rows(int kx, int ky) {
    int h = BEAM_LDS / (16 * (BT - 1 + ky) + 4 * ky) - (BT - 1);
    if (h > kx) h = kx;
    return h < 1 ? 1 : h;
}

__global__ __launch_bounds__(256) void k_convolve_beam(const float* __restrict__ in,
                                                       float* __restrict__ out, int nx, int ny,
                                                       int nc, const float* __restrict__ K,
                                                       int kx, int ky, int band) {
    extern __shared__ float4 lds[];
    const int tiles_y = (ny + BT - 1) / BT;
    const int tile = blockIdx.x;
    const int x0 = (tile / tiles_y) * BT, y0 = (tile % tiles_y) * BT;
    const int c0 = blockIdx.y * 4;
    const int t = threadIdx.x;
    const int tx = t / BT, ty = t % BT;
    const int hx = kx / 2, hy = ky / 2;
    const int wy = BT - 1 + ky;  // halo width along y
    float4* halo = lds;
    float* taps = reinterpret_cast<float*>(lds + (BT - 1 + band) * wy);

    // each beam row is summed in FP32 and added to FP64 totals, so the FP32
    // rounding grows with ky, not with kx * ky
    double ax = 0.0, ay = 0.0, az = 0.0, aw = 0.0;
    for (int a0 = 0; a0 < kx; a0 += band) {
        const int H = min(band, kx - a0);
        // output row x0 + tx reads input rows x0 + tx + hx - a for a in [a0, a0 + H)
        const int rx0 = x0 + hx - a0 - (H - 1);  // input row of halo row 0
        const int ry0 = y0 + hy - (ky - 1);      // input column of halo column 0
        const int hrows = BT - 1 + H;
        __syncthreads();  // the previous band's halo is no longer read
        for (int e = t; e < hrows * wy; e += 256) {
            const int r = e / wy, q = e - r * wy;
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
            halo[e] = v;
        }
        for (int e = t; e < H * ky; e += 256) taps[e] = K[(size_t)a0 * ky + e];
        __syncthreads();
        // tap (a0 + da, b) reads halo row tx + H - 1 - da, column ty + ky - 1 - b
        for (int da = 0; da < H; ++da) {
            const float4* hrow = halo + (tx + H - 1 - da) * wy + ty + ky - 1;
            const float* krow = taps + da * ky;
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int b = 0; b < ky; ++b) {
                const float k = krow[b];
                const float4 v = hrow[-b];
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
    const int band = beam_band_rows(kx, ky);
    return sizeof(float4) * (size_t)(BT - 1 + band) * (BT - 1 + ky) +
           sizeof(float) * (size_t)band * ky;
}

hipError_t convolve_beam(const float* in, float* out, int nx, int ny, int nc, const float* K,
                         int kx, int ky, hipStream_t s) {
    const int band = beam_band_rows(kx, ky);
    const size_t lds = beam_lds_bytes(kx, ky);
    const unsigned tiles = (unsigned)(((nx + BT - 1) / BT) * ((ny + BT - 1) / BT));
    const unsigned cq = (unsigned)((nc + 3) / 4);
    hipLaunchKernelGGL(k_convolve_beam, dim3(tiles, cq), dim3(256), lds, s, in, out, nx, ny, nc, K,
                       kx, ky, band);
    return hipGetLastError();
}

}  // namespace mi
