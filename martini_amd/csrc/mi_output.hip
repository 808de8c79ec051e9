// The FITS data unit of a cube, made on the device: Martini.write_fits
// (martini/martini.py:939-1055) drops the pad, takes the array's transpose
// (:1049, so that FITS axis 1 -- x -- runs fastest, then y, then channel),
// casts it (:1050-1051) and has astropy write it big-endian.  Here one kernel
// reads the padded [nx + 2 padx][ny + 2 pady][nc] cube (FP32 as the gathers
// leave it, or FP64 as the host holds it), and writes the unpadded region in
// FITS order as big-endian float32 or float64, ready to follow the header in
// the file; the same pass finds DATAMIN and DATAMAX.
//
// A work group moves 64 (x) x 64 (channel) tiles of 8 rows y through LDS:
// read along channels (the cube's contiguous axis), written along x (the
// file's).  Min and max use order-preserving integer images of the values,
// reduced in the work group, so their atomics give the same result in any
// order.
#include <hip/hip_runtime.h>

#include "../../include/martini_mi355x.h"
#include "mi_internal.h"

namespace mi {

constexpr int FT = 64;  // tile edge (x and channel)
constexpr int YB = 8;   // rows y per work group

// Unsigned image of a double whose integer order is the numeric order.
__device__ inline uint64_t double_order(double f) {
    const uint64_t u = (uint64_t)__double_as_longlong(f);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ inline double order_double(uint64_t u) {
    return __longlong_as_double((long long)((u >> 63) ? (u & 0x7fffffffffffffffull) : ~u));
}

__global__ void k_range_init(unsigned long long* __restrict__ range) {
    if (threadIdx.x == 0) {
        range[0] = ~0ull;  // running min
        range[1] = 0ull;   // running max
    }
}

__global__ void k_range_finish(unsigned long long* __restrict__ range) {
    if (threadIdx.x < 2) {
        const double f = order_double(range[threadIdx.x]);
        reinterpret_cast<double*>(range)[threadIdx.x] = f;
    }
}

template <typename In, bool F64>
__global__ __launch_bounds__(256) void k_cube_to_fits(const In* __restrict__ cube, int nx, int ny,
                                                      int nc, int padx, int pady,
                                                      void* __restrict__ out,
                                                      unsigned long long* __restrict__ range) {
    __shared__ In tile[FT][FT + 1];
    __shared__ unsigned long long wlo[4], whi[4];
    const int xt = (nx + FT - 1) / FT;
    const int x0 = (blockIdx.x % xt) * FT, c0 = (blockIdx.x / xt) * FT;
    const int t = threadIdx.x;
    const size_t NY = (size_t)ny + 2 * (size_t)pady;
    const size_t plane = (size_t)nx * ny;
    unsigned long long lo = ~0ull, hi = 0ull;
    const int y1 = min(ny, (int)(blockIdx.y + 1) * YB);
    for (int y = blockIdx.y * YB; y < y1; ++y) {
        for (int e = t; e < FT * FT; e += 256) {
            const int xi = e / FT, ci = e % FT;
            const int x = x0 + xi, c = c0 + ci;
            In v = 0;
            if (x < nx && c < nc) {
                v = cube[((size_t)(x + padx) * NY + (size_t)(y + pady)) * nc + c];
                const unsigned long long o = double_order((double)v);
                lo = o < lo ? o : lo;
                hi = o > hi ? o : hi;
            }
            tile[xi][ci] = v;
        }
        __syncthreads();
        for (int e = t; e < FT * FT; e += 256) {
            const int ci = e / FT, xi = e % FT;
            const int x = x0 + xi, c = c0 + ci;
            if (x >= nx || c >= nc) continue;
            const size_t at = (size_t)c * plane + (size_t)y * nx + x;
            if (F64) {
                const double v = (double)tile[xi][ci];
                reinterpret_cast<uint64_t*>(out)[at] =
                    __builtin_bswap64((uint64_t)__double_as_longlong(v));
            } else {
                const float v = (float)tile[xi][ci];
                reinterpret_cast<uint32_t*>(out)[at] = __builtin_bswap32(__float_as_uint(v));
            }
        }
        __syncthreads();  // the tile is rewritten for the next row
    }
    // block reduction, then one atomic pair per work group
    for (int d = 32; d > 0; d >>= 1) {
        const unsigned long long ol = __shfl_xor(lo, d, 64), oh = __shfl_xor(hi, d, 64);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
    }
    if ((t & 63) == 0) {
        wlo[t >> 6] = lo;
        whi[t >> 6] = hi;
    }
    __syncthreads();
    if (t == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = wlo[w] < lo ? wlo[w] : lo;
            hi = whi[w] > hi ? whi[w] : hi;
        }
        atomicMin(&range[0], lo);
        atomicMax(&range[1], hi);
    }
}

template <typename In>
static void launch_to_fits(const In* cube, int nx, int ny, int nc, int padx, int pady, bool f64,
                           void* out, unsigned long long* range, hipStream_t s) {
    const dim3 grid((unsigned)(((nx + FT - 1) / FT) * ((nc + FT - 1) / FT)),
                    (unsigned)((ny + YB - 1) / YB));
    if (f64)
        hipLaunchKernelGGL((k_cube_to_fits<In, true>), grid, dim3(256), 0, s, cube, nx, ny, nc,
                           padx, pady, out, range);
    else
        hipLaunchKernelGGL((k_cube_to_fits<In, false>), grid, dim3(256), 0, s, cube, nx, ny, nc,
                           padx, pady, out, range);
}

hipError_t cube_to_fits(const void* cube, bool cube_f64, int nx, int ny, int nc, int padx,
                        int pady, bool f64, void* out, double* range, hipStream_t s) {
    auto* r = reinterpret_cast<unsigned long long*>(range);
    hipLaunchKernelGGL(k_range_init, dim3(1), dim3(64), 0, s, r);
    if (cube_f64)
        launch_to_fits(static_cast<const double*>(cube), nx, ny, nc, padx, pady, f64, out, r, s);
    else
        launch_to_fits(static_cast<const float*>(cube), nx, ny, nc, padx, pady, f64, out, r, s);
    hipLaunchKernelGGL(k_range_finish, dim3(1), dim3(64), 0, s, r);
    return hipGetLastError();
}

}  // namespace mi
