// FP64 setup on the device: raw particle arrays (positions, velocities,
// masses, smoothing lengths, temperatures) -> mi_particle records, pruned and
// compacted in input order.  The same sequence the host runs in
// martini_amd (sources.py / datacube.py / sph_kernels.py / martini.py /
// _pack.py), restating:
//   SPHSource.rotate and the observed frame     sph_source.py:276-296, :442-526
//   sky coordinates                              sph_source.py:265-326
//   TAN + linear spectral WCS                    datacube.py:426-486, sph_source.py:328-362
//   smoothing lengths, adaptive choice, ranges   sph_kernels.py:235-262, :1241-1274
//   pruning                                      martini.py:168-241
//   thermal line width                           spectral_models.py:465-485
// Everything is FP64; the records carry the same integer/fraction splits as
// the host packer.
//
// Two passes over the particles (keep flags, then the records at their scan
// offsets) recompute the setup instead of storing it: the arithmetic is far
// cheaper than 48 more bytes per particle of HBM traffic.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>

#include "../../include/martini_mi355x.h"
#include "mi_device.h"
#include "mi_internal.h"

namespace mi {

constexpr double PI = 3.14159265358979323846;
constexpr double DEG_PER_RAD = 180.0 / PI;
constexpr double RAD_PER_DEG = PI / 180.0;
constexpr double ARCSEC_PER_RAD = 180.0 / PI * 3600.0;
constexpr double K_B = 1.380649e-23;       // J/K
constexpr double M_P = 1.67262192369e-27;  // kg
constexpr double HI_FREQ = 1.420405751e9;  // Hz
constexpr double C_LIGHT = 299792458.0;    // m/s
constexpr double MHI_TO_JY = 2.36e5;
constexpr double WINDOW_ARG = 4.0;         // erf argument of the spectral window edge
constexpr double MAX_INDEX = 2097152.0;    // 2^21: |pixel coordinate| bound
constexpr double MAX_RANGE = 4194304.0;    // 2^22: r bound (r_kind has 24 bits for it)

struct Setup {
    bool keep;
    uint32_t flags;
    uint64_t mask_px;  // pixels of the reference's square mask inside the padded cube
    mi_particle rec;
};

__device__ inline void matvec(const double* m, const double* v, double* out) {
    out[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    out[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    out[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

__device__ inline double pymod360(double x) { return x - 360.0 * floor(x / 360.0); }

__device__ Setup setup_particle(int64_t i, const double* __restrict__ xyz,
                                const double* __restrict__ vxyz, const double* __restrict__ mHI,
                                const double* __restrict__ hsm, const double* __restrict__ T,
                                const double* __restrict__ edges, const mi_observation& o) {
    Setup S;
    S.flags = 0;
    // source frame: the source's rotation, then the observed frame
    double x0[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    double v0[3] = {vxyz[3 * i], vxyz[3 * i + 1], vxyz[3 * i + 2]};
    double xr[3], vr[3];
    matvec(o.rotation, x0, xr);
    matvec(o.rotation, v0, vr);
    const double xk[3] = {xr[0] / 1000.0, xr[1] / 1000.0, xr[2] / 1000.0};
    double pos[3], vel[3];
    matvec(o.sky_rotation, xk, pos);
    matvec(o.sky_rotation, vr, vel);
    const double hub = o.hubble_h * 100.0;
    for (int a = 0; a < 3; ++a) {
        pos[a] += o.distance_mpc * o.pointing[a];
        vel[a] += hub * pos[a] + o.vpeculiar_kms * o.pointing[a];
    }
    const double dist = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
    const double vlos = (pos[0] * vel[0] + pos[1] * vel[1] + pos[2] * vel[2]) / dist;
    const double ra_deg = pymod360(atan2(pos[1], pos[0]) * DEG_PER_RAD);
    const double dec_deg = atan2(pos[2], hypot(pos[0], pos[1])) * DEG_PER_RAD;

    // TAN projection about the cube centre, RA to the left
    const double a0 = o.cube_ra_deg * RAD_PER_DEG, d0 = o.cube_dec_deg * RAD_PER_DEG;
    const double ar = ra_deg * RAD_PER_DEG, dr = dec_deg * RAD_PER_DEG;
    const double cos_d = cos(dr), sin_d = sin(dr);
    const double cos_da = cos(ar - a0);
    const double den = sin_d * sin(d0) + cos_d * cos(d0) * cos_da;
    const double nan = __builtin_nan("");
    double x = nan, y = nan;
    if (den > 0.0) {
        const double xi = cos_d * sin(ar - a0) / den * DEG_PER_RAD;
        const double eta = (sin_d * cos(d0) - cos_d * sin(d0) * cos_da) / den * DEG_PER_RAD;
        x = o.crpix[0] - 1.0 + xi / o.cdelt[0];
        y = o.crpix[1] - 1.0 + eta / o.cdelt[1];
    }
    const double v_ms = vlos * 1e3;
    const double spec = o.freq_channels ? HI_FREQ * (1.0 - v_ms / C_LIGHT) : v_ms;
    const double z = o.crpix[2] - 1.0 + (spec - o.crval2) / o.cdelt[2];

    // smoothing length in pixels and the kernel tier
    const double h_px = atan(hsm[i] / (dist * 1000.0)) * ARCSEC_PER_RAD / o.px_size_arcsec;
    int use = 0;
    if (o.n_tiers > 1) {
        int idx = -1;
        for (int k = 0; k < o.n_tiers && idx < 0; ++k) {
            const double he = h_px * o.tier_rescale[k];
            if (he >= o.tier_min[k] && he <= o.tier_max[k]) idx = k;
        }
        if (idx < 0) S.flags |= MI_PREP_INVALID_KERNEL;
        use = idx < 0 ? 0 : idx;
    } else if (!(h_px >= o.tier_min[0] && h_px <= o.tier_max[0])) {
        S.flags |= MI_PREP_INVALID_KERNEL;  // a fixed kernel validates the raw h_px
    }
    const double r = ceil(h_px * o.tier_size[use]);
    const double h_eff = h_px * o.tier_rescale[use];
    const int kind = o.tier_kind[use];

    // line width (km/s) and pruning
    const double m = mHI[i];
    double sigma = 0.0;
    if (o.spectrum == MI_SPECTRUM_GAUSS)
        sigma = o.sigma_kms >= 0.0 ? o.sigma_kms : sqrt(K_B * T[i] / M_P) / 1000.0;
    bool reject = false;
    if (o.prune) {
        const double hw = sigma / o.dv_max_kms;
        const int nxp = o.nx + 2 * o.padx, nyp = o.ny + 2 * o.pady;
        reject = (x + r < 0.0) || (y + r < 0.0) || (x - r > (double)nxp) || (y - r > (double)nyp) ||
                 isnan(x) || isnan(y) || (z + 4.0 * hw < 0.0) || (z - 4.0 * hw > (double)o.nc) ||
                 m == 0.0;
    }
    S.keep = !reject;
    S.mask_px = 0;
    if (!S.keep) {
        S.flags = 0;
        return S;
    }

    // record: footprint fields (_pack.kernel_fields), padded extent
    const int nx = o.nx + 2 * o.padx, ny = o.ny + 2 * o.pady;
    const bool alive = isfinite(x) && isfinite(y) && r >= 0.0 && x + r >= 0.0 && y + r >= 0.0 &&
                       x - r <= (double)(nx - 1) && y - r <= (double)(ny - 1);
    if (alive) {  // |i - x| <= r, |j - y| <= r (martini.py:271-274) within the padded cube
        const double mx = fmin((double)(nx - 1), floor(x + r)) - fmax(0.0, ceil(x - r)) + 1.0;
        const double my = fmin((double)(ny - 1), floor(y + r)) - fmax(0.0, ceil(y - r)) + 1.0;
        if (mx > 0.0 && my > 0.0) S.mask_px = (uint64_t)mx * (uint64_t)my;
    }
    const double xs = alive ? x : 0.0, ys = alive ? y : 0.0;
    if (fabs(xs) >= MAX_INDEX || fabs(ys) >= MAX_INDEX) S.flags |= MI_PREP_COORD_LIMIT;
    double ix = floor(xs), iy = floor(ys);
    double fx = xs - ix, fy = ys - iy;
    double rr = alive ? fmin(r, MAX_RANGE) : -1.0;
    if (kind == MI_KERNEL_DIRAC) {
        const double pi_ = floor(xs + 0.5), pj = floor(ys + 0.5);
        const bool ok = fabs(xs - pi_) < 0.5 && fabs(ys - pj) < 0.5 && fabs(xs - pi_) <= rr &&
                        fabs(ys - pj) <= rr && pi_ >= 0.0 && pi_ <= (double)(nx - 1) && pj >= 0.0 &&
                        pj <= (double)(ny - 1) && alive;
        ix = pi_;
        iy = pj;
        fx = 0.0;
        fy = 0.0;
        rr = ok ? 0.0 : -1.0;
    }
    mi_particle& p = S.rec;
    p.ix = (int32_t)ix;
    p.iy = (int32_t)iy;
    p.fx = (float)fx;
    p.fy = (float)fy;
    p.h = (float)h_eff;
    p.r_kind = (int32_t)((int64_t)rr * 256 + (int64_t)o.tier_gidx[use] * 16 + kind);

    // spectral fields (_pack.spectral_fields)
    const int nc = o.nc;
    const double u = (v_ms - o.e0_ms) / o.de_ms;
    const bool finite = isfinite(u);
    const double uu = finite ? u : 0.0;
    const double k0 = floor(uu);
    const double fu = uu - k0;
    const double ade = fabs(o.de_ms);
    const double amp = m * pow(dist, -2.0) / (ade / 1e3) / MHI_TO_JY * o.out_scale;
    double c_lo, c_hi, s = 0.0;
    if (o.spectrum == MI_SPECTRUM_GAUSS) {
        s = ade / (1.4142135623730951 * sigma * 1e3);
        const double half = WINDOW_ARG / s;
        c_lo = floor(uu - 1.0 - half) + 1.0;
        c_hi = ceil(uu + half) - 1.0;
    } else {
        c_lo = INFINITY;
        c_hi = -INFINITY;
        for (int dd = -1; dd <= 1; ++dd) {
            const double cf = k0 + dd;
            const int c = (int)fmin(fmax(cf, 0.0), (double)(nc - 1));
            const double lo_e = fmin(edges[c], edges[c + 1]), hi_e = fmax(edges[c], edges[c + 1]);
            if (v_ms >= lo_e && v_ms <= hi_e && finite && cf >= 0.0 && cf <= (double)(nc - 1)) {
                c_lo = fmin(c_lo, (double)c);
                c_hi = fmax(c_hi, (double)c);
            }
        }
    }
    c_lo = fmin(fmax(finite ? c_lo : (double)nc, 0.0), (double)nc);
    c_hi = fmin(fmax(finite ? c_hi : -1.0, -1.0), (double)(nc - 1));
    if (c_lo > c_hi) {
        c_lo = 1.0;
        c_hi = 0.0;
    }
    // a shared cube addresses this source's channels from channel_offset on
    const bool empty = c_lo > c_hi;
    p.c_lo = (int32_t)c_lo + (empty ? 0 : o.channel_offset);
    p.c_hi = (int32_t)c_hi + (empty ? 0 : o.channel_offset);
    p.k0 = (int32_t)fmin(fmax(k0, -1073741824.0), 1073741824.0) + o.channel_offset;
    p.fu = (float)fu;
    p.s = (float)s;
    p.amp = (float)amp;
    return S;
}

__global__ void k_prep_init(unsigned long long* counts) {
    if (threadIdx.x < 4) counts[threadIdx.x] = 0ull;
}

__global__ void k_prep_flags(const double* __restrict__ xyz, const double* __restrict__ vxyz,
                             const double* __restrict__ mHI, const double* __restrict__ hsm,
                             const double* __restrict__ T, const double* __restrict__ edges,
                             int64_t n, mi_observation o, uint32_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keep[i] = setup_particle(i, xyz, vxyz, mHI, hsm, T, edges, o).keep ? 1u : 0u;
}

__global__ void k_prep_write(const double* __restrict__ xyz, const double* __restrict__ vxyz,
                             const double* __restrict__ mHI, const double* __restrict__ hsm,
                             const double* __restrict__ T, const double* __restrict__ edges,
                             int64_t n, mi_observation o, Grid g, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ off, mi_particle* __restrict__ out,
                             unsigned long long* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long pairs = 0, px = 0;
    uint32_t flags = 0;
    if (i < n && keep[i]) {
        const Setup S = setup_particle(i, xyz, vxyz, mHI, hsm, T, edges, o);
        out[off[i]] = S.rec;
        int a, b, c, d, e, f;
        pairs = cells_of(S.rec, g, a, b, c, d, e, f);
        flags = S.flags;
        px = S.mask_px;
    }
    // one atomic per wave (integer sums: order-independent)
#pragma unroll
    for (int dd = 32; dd > 0; dd >>= 1) {
        pairs += __shfl_xor(pairs, dd, 64);
        px += __shfl_xor(px, dd, 64);
        flags |= __shfl_xor(flags, dd, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (pairs) atomicAdd(&counts[1], pairs);
        if (flags) atomicOr(&counts[2], (unsigned long long)flags);
        if (px) atomicAdd(&counts[3], px);
    }
    if (i == n - 1) atomicAdd(&counts[0], (unsigned long long)off[i] + keep[i]);
}

// Workspace: keep flags and their exclusive scan (uint32[n] each), then the
// scan's temporary storage.

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

hipError_t prep_layout(int64_t n, PrepLayout* L) {
    size_t tb = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, tb, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           0u, (size_t)(n > 0 ? n : 1), rocprim::plus<uint32_t>(), 0);
    if (e != hipSuccess) return e;
    size_t o = 0;
    L->keep = o;
    o += align256(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    L->off = o;
    o += align256(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    L->temp = o;
    L->temp_bytes = tb;
    o += align256(tb);
    L->total = o;
    return hipSuccess;
}

hipError_t prepare_particles(const double* xyz, const double* vxyz, const double* mHI,
                             const double* hsm, const double* T, const double* edges, int64_t n,
                             const mi_observation& o, const Grid& g, mi_particle* out,
                             unsigned long long* counts, char* ws, const PrepLayout& L,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_prep_init, dim3(1), dim3(64), 0, s, counts);
    if (n == 0) return hipGetLastError();
    const int T_ = 256;
    const unsigned blocks = (unsigned)((n + T_ - 1) / T_);
    uint32_t* keep = (uint32_t*)(ws + L.keep);
    uint32_t* off = (uint32_t*)(ws + L.off);
    hipLaunchKernelGGL(k_prep_flags, dim3(blocks), dim3(T_), 0, s, xyz, vxyz, mHI, hsm, T, edges,
                       n, o, keep);
    size_t tb = L.temp_bytes;
    hipError_t e = rocprim::exclusive_scan(ws + L.temp, tb, keep, off, 0u, (size_t)n,
                                           rocprim::plus<uint32_t>(), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_prep_write, dim3(blocks), dim3(T_), 0, s, xyz, vxyz, mHI, hsm, T, edges,
                       n, o, g, keep, off, out, counts);
    return hipGetLastError();
}

}  // namespace mi
