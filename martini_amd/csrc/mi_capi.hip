// extern "C" entry points of libmartini_mi355x (declared in
// include/martini_mi355x.h).  Argument checking, workspace layout, and the
// stage sequence bin -> gather -> reduce.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/martini_mi355x.h"
#include "mi_device.h"
#include "mi_internal.h"

using namespace mi;

// Workspace of mi_insert_noalloc: a static device array, one per GPU, held
// until the process ends.  Host code never passes this name to a kernel or a
// copy: it uses the device address that hipGetSymbolAddress returns for the
// current device (scratch_address).
static __device__ __align__(256) unsigned char scratch[MI_SCRATCH_BYTES];

// The device address of `scratch` on the current device, resolved once per
// device with that device current.
static hipError_t scratch_address(void** out) {
    static std::mutex mu;
    static void* ptr[MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (p == nullptr) return hipErrorInvalidSymbol;
        ptr[dev] = p;
    }
    *out = ptr[dev];
    return hipSuccess;
}

static bool desc_ok(const mi_cube_desc* d) {
    if (!d) return false;
    if (d->nx <= 0 || d->ny <= 0 || d->nc <= 0) return false;
    if (d->spectrum != MI_SPECTRUM_GAUSS && d->spectrum != MI_SPECTRUM_DIRAC &&
        d->spectrum != MI_SPECTRUM_TABLE)
        return false;
    if (d->spectrum == MI_SPECTRUM_TABLE && (!d->spectra || d->spectra_ld < d->nc)) return false;
    if (d->weights_n < 0 || (d->weights_n > 0 && !d->weights)) return false;
    return true;
}

// Limits that the key and index widths impose.
static int check_limits(const Grid& g, int64_t n, int64_t n_pairs) {
    if (g.ncells >= ((int64_t)1 << 21)) return MI_ERR_LIMIT;  // cell id fits 22 key bits
    if (n >= ((int64_t)1 << 31) || n_pairs >= ((int64_t)1 << 31)) return MI_ERR_LIMIT;
    if (g.nx > 32767 || g.ny > 32767) return MI_ERR_LIMIT;
    return MI_OK;
}

extern "C" {

int mi_abi_version(void) { return MI_ABI_VERSION; }

const char* mi_strerror(int code) {
    switch (code) {
    case MI_OK: return "success";
    case MI_ERR_ARGUMENT: return "invalid argument";
    case MI_ERR_WORKSPACE: return "workspace too small for this input";
    case MI_ERR_DEVICE: return "HIP runtime error";
    case MI_ERR_UNSUPPORTED: return "unsupported kernel, spectrum or method";
    case MI_ERR_LIMIT: return "input exceeds a size limit of the library";
    case MI_ERR_SCRATCH:
        return "input needs more workspace than the static scratch of mi_insert_noalloc "
               "(MI_SCRATCH_BYTES)";
    }
    return "unknown error";
}

int64_t mi_count_pairs_host(const mi_particle* h_particles, int64_t n, const mi_cube_desc* desc) {
    if (!desc_ok(desc) || n < 0 || (n > 0 && !h_particles)) return MI_ERR_ARGUMENT;
    Grid g = make_grid(*desc);
    int64_t tot = 0;
    for (int64_t k = 0; k < n; ++k) {
        int a, b, c, d, e, f;
        tot += cells_of(h_particles[k], g, a, b, c, d, e, f);
    }
    return tot;
}

int mi_insert_workspace_size(const mi_cube_desc* desc, int64_t n, int64_t n_pairs,
                             size_t* bytes) {
    if (!desc_ok(desc) || n < 0 || n_pairs < 0 || !bytes) return MI_ERR_ARGUMENT;
    Grid g = make_grid(*desc);
    int rc = check_limits(g, n, n_pairs);
    if (rc) return rc;
    Layout L;
    if (layout_for(g, n, n_pairs, &L) != hipSuccess) return MI_ERR_DEVICE;
    *bytes = L.total;
    return MI_OK;
}

static int prepare(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
                   const mi_cube_desc* desc, const void* d_workspace, size_t workspace_bytes,
                   Grid* g, Layout* L) {
    if (!desc_ok(desc) || n < 0 || n_pairs < 0 || !d_workspace) return MI_ERR_ARGUMENT;
    if (n > 0 && !d_particles) return MI_ERR_ARGUMENT;
    if (reinterpret_cast<uintptr_t>(d_workspace) & 255u) return MI_ERR_ARGUMENT;
    *g = make_grid(*desc);
    int rc = check_limits(*g, n, n_pairs);
    if (rc) return rc;
    if (layout_for(*g, n, n_pairs, L) != hipSuccess) return MI_ERR_DEVICE;
    if (workspace_bytes < L->total) return MI_ERR_WORKSPACE;
    return MI_OK;
}

int mi_bin(const mi_particle* d_particles, int64_t n, int64_t n_pairs, const mi_cube_desc* desc,
           void* d_workspace, size_t workspace_bytes, int method, void* stream) {
    if (method != MI_METHOD_PLAIN_F32 && method != MI_METHOD_MFMA_BF16X3) return MI_ERR_UNSUPPORTED;
    Grid g;
    Layout L;
    int rc = prepare(d_particles, n, n_pairs, desc, d_workspace, workspace_bytes, &g, &L);
    if (rc) return rc;
    if (bin_particles(d_particles, n, n_pairs, g, *desc, (char*)d_workspace, L,
                      method == MI_METHOD_MFMA_BF16X3, (hipStream_t)stream) != hipSuccess)
        return MI_ERR_DEVICE;
    return MI_OK;
}

int mi_gather(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
              const mi_cube_desc* desc, float* d_cube, void* d_workspace, size_t workspace_bytes,
              int method, void* stream) {
    if (!d_cube) return MI_ERR_ARGUMENT;
    if (method != MI_METHOD_PLAIN_F32 && method != MI_METHOD_MFMA_BF16X3) return MI_ERR_UNSUPPORTED;
    Grid g;
    Layout L;
    int rc = prepare(d_particles, n, n_pairs, desc, d_workspace, workspace_bytes, &g, &L);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)d_workspace;
    hipError_t e = method == MI_METHOD_PLAIN_F32 ? gather_plain(d_particles, g, *desc, ws, L, d_cube, s)
                                                 : gather_mfma(d_particles, g, *desc, ws, L, d_cube, s);
    if (e == hipErrorNotSupported) return MI_ERR_UNSUPPORTED;
    if (e != hipSuccess) return MI_ERR_DEVICE;
    if (reduce_partials(g, ws, L, d_cube, s) != hipSuccess) return MI_ERR_DEVICE;
    return MI_OK;
}

int mi_status(const mi_cube_desc* desc, int64_t n, int64_t n_pairs, const void* d_workspace,
              uint32_t out[4], void* stream) {
    if (!out) return MI_ERR_ARGUMENT;
    Grid g;
    Layout L;
    if (!desc_ok(desc) || !d_workspace || n < 0 || n_pairs < 0) return MI_ERR_ARGUMENT;
    if (reinterpret_cast<uintptr_t>(d_workspace) & 255u) return MI_ERR_ARGUMENT;
    g = make_grid(*desc);
    if (layout_for(g, n, n_pairs, &L) != hipSuccess) return MI_ERR_DEVICE;
    uint32_t st[8];
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(st, (const char*)d_workspace + L.status, sizeof(st), hipMemcpyDeviceToHost,
                       s) != hipSuccess)
        return MI_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MI_ERR_DEVICE;
    out[0] = st[ST_PAIRS];
    out[1] = st[ST_OVERFLOW];
    out[2] = st[ST_ITEMS];
    out[3] = st[ST_MULTI];
    if (st[ST_OVERFLOW]) return MI_ERR_WORKSPACE;
    return st[ST_BADREC] ? MI_ERR_ARGUMENT : MI_OK;
}

int mi_insert(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
              const mi_cube_desc* desc, float* d_cube, void* d_workspace, size_t workspace_bytes,
              int method, void* stream) {
    if (!d_cube) return MI_ERR_ARGUMENT;
    if (method != MI_METHOD_PLAIN_F32 && method != MI_METHOD_MFMA_BF16X3) return MI_ERR_UNSUPPORTED;
    int rc = mi_bin(d_particles, n, n_pairs, desc, d_workspace, workspace_bytes, method, stream);
    if (rc) return rc;
    rc = mi_gather(d_particles, n, n_pairs, desc, d_cube, d_workspace, workspace_bytes, method,
                   stream);
    if (rc) return rc;
    uint32_t st[4];
    return mi_status(desc, n, n_pairs, d_workspace, st, stream);
}

int mi_insert_noalloc(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
                      const mi_cube_desc* desc, float* d_cube, int method, void* stream) {
    if (!d_cube) return MI_ERR_ARGUMENT;
    if (method != MI_METHOD_PLAIN_F32 && method != MI_METHOD_MFMA_BF16X3) return MI_ERR_UNSUPPORTED;
    size_t need = 0;
    int rc = mi_insert_workspace_size(desc, n, n_pairs, &need);
    if (rc) return rc;
    if (need > MI_SCRATCH_BYTES) return MI_ERR_SCRATCH;
    void* ws = nullptr;
    if (scratch_address(&ws) != hipSuccess) return MI_ERR_DEVICE;
    return mi_insert(d_particles, n, n_pairs, desc, d_cube, ws, MI_SCRATCH_BYTES, method, stream);
}

int mi_kernel_weights(int kind, float gauss_t, float gauss_inv_norm, const float* d_dx,
                      const float* d_dy, const float* d_h, float* d_out, int64_t n,
                      void* stream) {
    if (n < 0 || (n > 0 && (!d_dx || !d_dy || !d_h || !d_out))) return MI_ERR_ARGUMENT;
    if (kind < MI_KERNEL_CUBIC || kind > MI_KERNEL_DIRAC) return MI_ERR_UNSUPPORTED;
    if (launch_weights(kind, gauss_t, gauss_inv_norm, d_dx, d_dy, d_h, d_out, n,
                       (hipStream_t)stream) != hipSuccess)
        return MI_ERR_DEVICE;
    return MI_OK;
}

int mi_spectra(const mi_particle* d_particles, int64_t n, const mi_cube_desc* desc, float* d_out,
               void* stream) {
    if (!desc_ok(desc) || n < 0 || (n > 0 && (!d_particles || !d_out))) return MI_ERR_ARGUMENT;
    if (launch_spectra(d_particles, n, *desc, d_out, (hipStream_t)stream) != hipSuccess)
        return MI_ERR_DEVICE;
    return MI_OK;
}

int mi_prepare_workspace_size(int64_t n, size_t* bytes) {
    if (n < 0 || !bytes) return MI_ERR_ARGUMENT;
    if (n >= ((int64_t)1 << 31)) return MI_ERR_LIMIT;
    PrepLayout L;
    if (prep_layout(n, &L) != hipSuccess) return MI_ERR_DEVICE;
    *bytes = L.total;
    return MI_OK;
}

static bool observation_ok(const mi_observation* o) {
    if (!o) return false;
    if (o->nx <= 0 || o->ny <= 0 || o->nc <= 0 || o->padx < 0 || o->pady < 0) return false;
    if (o->spectrum != MI_SPECTRUM_GAUSS && o->spectrum != MI_SPECTRUM_DIRAC) return false;
    if (o->n_tiers < 1 || o->n_tiers > 3) return false;
    for (int k = 0; k < o->n_tiers; ++k) {
        if (o->tier_kind[k] < MI_KERNEL_CUBIC || o->tier_kind[k] > MI_KERNEL_DIRAC) return false;
        if (o->tier_gidx[k] < 0 || o->tier_gidx[k] > 3) return false;
    }
    if (!(o->px_size_arcsec > 0.0) || !(o->dv_max_kms > 0.0) || o->de_ms == 0.0) return false;
    const int nct = o->nc_total > 0 ? o->nc_total : o->nc;
    if (o->channel_offset < 0 || o->channel_offset + o->nc > nct) return false;
    return true;
}

int mi_prepare(const double* d_xyz, const double* d_vxyz, const double* d_mHI,
               const double* d_hsm, const double* d_T, const double* d_edges, int64_t n,
               const mi_observation* obs, mi_particle* d_out, uint64_t* d_counts,
               void* d_workspace, size_t workspace_bytes, void* stream) {
    // table spectra come from the caller's host evaluation, not from raw arrays
    if (obs && obs->spectrum == MI_SPECTRUM_TABLE) return MI_ERR_UNSUPPORTED;
    if (!observation_ok(obs) || n < 0 || !d_counts) return MI_ERR_ARGUMENT;
    if (n > 0 && (!d_xyz || !d_vxyz || !d_mHI || !d_hsm || !d_out || !d_workspace))
        return MI_ERR_ARGUMENT;
    if (reinterpret_cast<uintptr_t>(d_workspace) & 255u) return MI_ERR_ARGUMENT;
    if (n > 0 && obs->spectrum == MI_SPECTRUM_GAUSS && obs->sigma_kms < 0.0 && !d_T)
        return MI_ERR_ARGUMENT;
    if (n > 0 && obs->spectrum == MI_SPECTRUM_DIRAC && !d_edges) return MI_ERR_ARGUMENT;
    if (n >= ((int64_t)1 << 31)) return MI_ERR_LIMIT;
    mi_cube_desc d{};
    d.nx = obs->nx + 2 * obs->padx;
    d.ny = obs->ny + 2 * obs->pady;
    d.nc = obs->nc_total > 0 ? obs->nc_total : obs->nc;
    d.spectrum = obs->spectrum;
    Grid g = make_grid(d);
    if (check_limits(g, n, 0) != MI_OK) return MI_ERR_LIMIT;
    PrepLayout L;
    if (prep_layout(n, &L) != hipSuccess) return MI_ERR_DEVICE;
    if (workspace_bytes < L.total) return MI_ERR_WORKSPACE;
    if (prepare_particles(d_xyz, d_vxyz, d_mHI, d_hsm, d_T, d_edges, n, *obs, g, d_out,
                          (unsigned long long*)d_counts, (char*)d_workspace, L,
                          (hipStream_t)stream) != hipSuccess)
        return MI_ERR_DEVICE;
    return MI_OK;
}

int mi_convolve_beam(const float* d_in, float* d_out, int32_t nx, int32_t ny, int32_t nc,
                     const float* d_kernel, int32_t kx, int32_t ky, void* stream) {
    if (!d_in || !d_out || !d_kernel || d_in == d_out) return MI_ERR_ARGUMENT;
    if (nx <= 0 || ny <= 0 || nc <= 0 || kx <= 0 || ky <= 0) return MI_ERR_ARGUMENT;
    if ((kx & 1) == 0 || (ky & 1) == 0) return MI_ERR_ARGUMENT;
    if (nx > 32767 || ny > 32767 || kx > 32767 || ky > 32767) return MI_ERR_LIMIT;
    if (convolve_beam(d_in, d_out, nx, ny, nc, d_kernel, kx, ky, (hipStream_t)stream) !=
        hipSuccess)
        return MI_ERR_DEVICE;
    return MI_OK;
}

int mi_cube_to_fits(const void* d_cube, int32_t cube_bits, int32_t nx, int32_t ny, int32_t nc,
                    int32_t padx, int32_t pady, int32_t bitpix, void* d_out, double* d_range,
                    void* stream) {
    if (!d_cube || !d_out || !d_range) return MI_ERR_ARGUMENT;
    if (cube_bits != 32 && cube_bits != 64) return MI_ERR_ARGUMENT;
    if (nx <= 0 || ny <= 0 || nc <= 0 || padx < 0 || pady < 0) return MI_ERR_ARGUMENT;
    if (bitpix != -32 && bitpix != -64) return MI_ERR_ARGUMENT;
    if (nx > 32767 || ny > 32767 || padx > 32767 || pady > 32767) return MI_ERR_LIMIT;
    if (cube_to_fits(d_cube, cube_bits == 64, nx, ny, nc, padx, pady, bitpix == -64, d_out,
                     d_range, (hipStream_t)stream) != hipSuccess)
        return MI_ERR_DEVICE;
    return MI_OK;
}

}  // extern "C"
