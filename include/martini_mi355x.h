/*
 * martini_mi355x.h -- C ABI of the MI355X source-insertion library.
 *
 * Replaces the hot loop of Martini.insert_source_in_cube
 * (martini/martini.py:285-407, _evaluate_pixel_spectrum :243-283) together
 * with the per-particle kernel integrals (martini/sph_kernels.py:85-119 and
 * the concrete _kernel_integral methods) and the spectra
 * (martini/spectral_models.py:63-200, :347-425, :510-570).
 *
 * Conventions
 *   - All pointers named d_* are device pointers (HBM); all calls are
 *     asynchronous on the given stream unless documented otherwise.
 *   - Return value: 0 on success, a negative MI_ERR_* code on failure.
 *     mi_strerror() turns it into text.  No exception crosses the ABI.
 *   - Nothing is allocated inside the calls: the caller passes the
 *     workspace, sized by mi_insert_workspace_size().  The kernel weight
 *     tables are constant data in a static device array; the first call on a
 *     device that reads them (mi_insert, mi_gather, mi_kernel_weights) fills
 *     them on its stream and synchronises that stream once.
 *   - Calls are reentrant for distinct (workspace, stream) pairs.
 *   - The cube is FP32, row-major [nx][ny][nc] over the padded extent, and
 *     is ACCUMULATED (+=), as the reference accumulates into DataCube._array
 *     (martini/martini.py:338-362).
 */
#ifndef MARTINI_MI355X_H
#define MARTINI_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_ABI_VERSION 3

/* Error codes. */
#define MI_OK 0
#define MI_ERR_ARGUMENT (-1)   /* null pointer, negative size, bad descriptor */
#define MI_ERR_WORKSPACE (-2)  /* workspace too small for this input */
#define MI_ERR_DEVICE (-3)     /* a HIP runtime call failed */
#define MI_ERR_UNSUPPORTED (-4) /* kernel/spectrum kind not handled */
#define MI_ERR_LIMIT (-5)      /* input exceeds a documented size limit */
#define MI_ERR_SCRATCH (-6)    /* input needs more than MI_SCRATCH_BYTES (mi_insert_noalloc) */

/* Bytes of the static per-device scratch that mi_insert_noalloc works in. */
#define MI_SCRATCH_BYTES ((size_t)256 << 20)

/* Kernel kinds (low 4 bits of mi_particle.r_kind).  Concrete kernels of
 * martini/sph_kernels.py: _CubicSplineKernel :724, _WendlandC2Kernel :340,
 * _WendlandC6Kernel :481, _QuarticSplineKernel :1402, _GaussianKernel :897,
 * DiracDeltaKernel :1083.  Adaptive kernels are resolved per particle on the
 * host (_AdaptiveKernel._init_sm_lengths :1241-1274). */
#define MI_KERNEL_CUBIC 0
#define MI_KERNEL_WC2 1
#define MI_KERNEL_WC6 2
#define MI_KERNEL_QUARTIC 3
#define MI_KERNEL_GAUSS 4
#define MI_KERNEL_DIRAC 5
/* Any other kernel (a user _BaseSPHKernel subclass, or a subclass with its own
 * _kernel_integral, sph_kernels.py:51-337): the caller evaluates its
 * _kernel_integral over each particle's masked pixels, as _px_weight does per
 * pixel (sph_kernels.py:85-119, called at martini.py:275-277), into a device
 * table of weights (mi_cube_desc.weights) that the insertion reads. */
#define MI_KERNEL_TABLE 6

/* Spectral models (martini/spectral_models.py): GaussianSpectrum :304,
 * DiracDeltaSpectrum :488. */
#define MI_SPECTRUM_GAUSS 0
#define MI_SPECTRUM_DIRAC 1
#define MI_SPECTRUM_TABLE 2  /* spectra evaluated by the caller, read from a device table */

/* Insertion methods. */
#define MI_METHOD_PLAIN_F32 0   /* FP32 VALU tile gather */
#define MI_METHOD_MFMA_BF16X3 1 /* BF16 split-precision MFMA contraction */

/*
 * One particle after the host's FP64 setup (projection, smoothing length,
 * adaptive kernel choice, pruning).  48 bytes, 16-byte aligned.
 *
 * Pixel position x = ix + fx, y = iy + fy (0-based, padded cube pixels;
 * SPHSource.pixcoords, martini/sources/sph_source.py:328-362).  Splitting off
 * the integer part keeps the FP32 offsets x - i exact to 2^-24 px.
 * Mask: |i - x| <= r and |j - y| <= r (martini/martini.py:271-274).
 * Dirac-kernel particles carry their single pixel as (ix, iy), fx = fy = 0,
 * r = 0, or r < 0 when no pixel receives weight.
 *
 * Spectrum: channel edges e_k = e0 + k*de (uniform in velocity; both VRAD and
 * radio-FREQ cubes are).  u = (v - e0)/de = k0 + fu.  Channel c receives
 * amp * 0.5*(erf(s*(u-c)) - erf(s*(u-c-1))) for a Gaussian line with
 * s = |de|/(sqrt(2) sigma), or amp for c in [c_lo, c_hi] for a Dirac line.
 * [c_lo, c_hi] is the window outside which the line is below 1e-10 of its
 * flux (Gaussian) or zero (Dirac).  With MI_SPECTRUM_TABLE (any spectral
 * model's _BaseSpectrum.spectral_function, martini/spectral_models.py:
 * 216-273, evaluated by the caller) channel c receives
 * amp * spectra[k0 * spectra_ld + c] for c in [c_lo, c_hi]; fu and s are unused.
 *
 * MI_KERNEL_TABLE particles: h holds, as the bits of a uint32, the offset in
 * mi_cube_desc.weights of the particle's pixel weights, row-major over its
 * footprint clipped to the cube, rows i_lo..i_hi by columns j_lo..j_hi (the
 * pixels the mask admits); the weight of pixel (i, j) is
 * weights[h + (i - i_lo) * (j_hi - j_lo + 1) + (j - j_lo)].
 */
typedef struct mi_particle {
    int32_t ix, iy;
    float fx, fy;
    float h;         /* effective kernel size h_px * rescale, pixels */
    int32_t r_kind;  /* (r << 8) | (gaussian table index << 4) | kernel kind */
    int32_t c_lo, c_hi;
    int32_t k0;
    float fu;
    float s;
    float amp;       /* mHI / d^2 / |dv| / 2.36e5 * output scale (Jy per unit weight) */
} mi_particle;

typedef struct mi_cube_desc {
    int32_t nx, ny, nc;      /* padded cube extent */
    int32_t spectrum;        /* MI_SPECTRUM_* */
    float gauss_t[4];        /* Gaussian-kernel truncations by table index */
    float gauss_inv_norm[4]; /* 1 / _GaussianKernel.norm by table index */
    const float* spectra;    /* MI_SPECTRUM_TABLE: device table of spectra, one row per
                                record's k0 (null otherwise) */
    int64_t spectra_ld;      /* its row stride in floats, >= nc */
    const float* weights;    /* MI_KERNEL_TABLE particles: device table of pixel weights
                                (null when no particle uses it) */
    int64_t weights_n;       /* its length in floats */
} mi_cube_desc;

/* Library ABI version (MI_ABI_VERSION). */
int mi_abi_version(void);

/* Text for an MI_ERR_* code. */
const char* mi_strerror(int code);

/*
 * Number of (particle, cell) pairs the binning of these particles produces:
 * a cell is an 8 x 16 pixel tile times a 128-channel block when the cube has
 * 128 channels or more (nc >= 128), else a 16 x 16 tile times a 64-channel
 * block.  Host-side count
 * from host records (no device access); the value sizes the workspace.
 */
int64_t mi_count_pairs_host(const mi_particle* h_particles, int64_t n,
                            const mi_cube_desc* desc);

/* Workspace bytes for mi_insert with n particles producing n_pairs pairs.
 * Every entry point that takes d_workspace requires it 256-byte aligned (as
 * hipMalloc returns it) and refuses it otherwise with MI_ERR_ARGUMENT. */
int mi_insert_workspace_size(const mi_cube_desc* desc, int64_t n, int64_t n_pairs,
                             size_t* bytes);

/*
 * Bin, sort and insert: cube += sum over particles of weight * spectrum.
 * Replaces _BaseMartini._insert_source_in_cube's pixel loop
 * (martini/martini.py:327-362).  d_cube must hold nx*ny*nc floats.
 * n_pairs must equal mi_count_pairs_host() for these particles (or bound it:
 * the call fails with MI_ERR_WORKSPACE, leaving the cube untouched, if the
 * device count exceeds it).  Synchronises the stream once, after binning.
 * method: MI_METHOD_*.
 */
int mi_insert(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
              const mi_cube_desc* desc, float* d_cube, void* d_workspace,
              size_t workspace_bytes, int method, void* stream);

/*
 * mi_insert without a caller workspace: the same binning, gather and status
 * check, working in a static device array of MI_SCRATCH_BYTES (256 MiB) that
 * the library holds on each GPU it runs on until the process ends.  Nothing is
 * allocated: no hipMalloc, no allocating library routine.  The caller passes
 * the particles and the cube as for mi_insert.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).
 * Returns MI_ERR_SCRATCH, before any launch, when this input's workspace
 * (mi_insert_workspace_size) exceeds MI_SCRATCH_BYTES.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (other threads,
 * other streams) race and must be serialised by the caller.  The first call on
 * a device resolves the scratch's device address and synchronises the stream
 * (as mi_insert does after binning).
 */
int mi_insert_noalloc(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
                      const mi_cube_desc* desc, float* d_cube, int method, void* stream);

/*
 * The two stages of mi_insert, asynchronous (no host synchronisation), for
 * callers that time or overlap them.  mi_bin sorts the particles into cells
 * inside the workspace; mi_gather then evaluates the contraction over those
 * cells with `method` and adds it to d_cube.  mi_bin's `method` names the
 * gather it prepares for: the MFMA form gets each cell's particles in record
 * order (records laid out by tile and pixel Z order make its 32-particle
 * chunks compact, and its skips are per chunk), the FP32 form an order by a
 * hash of the particle index (each chunk then spreads over the tile, which
 * balances its products over the waves).  Either order gives a correct gather
 * with either method.  The workspace must not be reused between the
 * two calls.  Overflow of n_pairs is reported by mi_status (which
 * synchronises the stream); a gather after an overflowed bin leaves the cube
 * untouched.
 */
int mi_bin(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
           const mi_cube_desc* desc, void* d_workspace, size_t workspace_bytes, int method,
           void* stream);
int mi_gather(const mi_particle* d_particles, int64_t n, int64_t n_pairs,
              const mi_cube_desc* desc, float* d_cube, void* d_workspace,
              size_t workspace_bytes, int method, void* stream);

/*
 * Synchronises the stream and reports the last binning in this workspace:
 * out[0] = pairs produced, out[1] = overflow flag, out[2] = work items,
 * out[3] = split work items (partial tiles).  Returns MI_ERR_WORKSPACE on
 * overflow, and MI_ERR_ARGUMENT when a record with a footprint has an
 * unknown kernel kind or MI_KERNEL_TABLE weights beyond mi_cube_desc.weights
 * (the binning then makes no work items: the cube is left untouched).
 */
int mi_status(const mi_cube_desc* desc, int64_t n, int64_t n_pairs, const void* d_workspace,
              uint32_t out[4], void* stream);

/*
 * Kernel integrals at n offsets: out[k] = W(dx[k], dy[k], h[k]) for one
 * concrete kernel kind (gaussian: truncation t, 1/norm).  Device arrays.
 * Replaces _BaseSPHKernel._px_weight / _kernel_integral
 * (martini/sph_kernels.py:85-119).
 */
int mi_kernel_weights(int kind, float gauss_t, float gauss_inv_norm, const float* d_dx,
                      const float* d_dy, const float* d_h, float* d_out, int64_t n,
                      void* stream);

/*
 * Spectra: out[p*nc + c] = amp_p * profile_p(c) for n particles (N x Nc,
 * FP32), from the same fields insertion uses.  Replaces
 * _BaseSpectrum.init_spectra's evaluation (martini/spectral_models.py:63-147).
 */
int mi_spectra(const mi_particle* d_particles, int64_t n, const mi_cube_desc* desc,
               float* d_out, void* stream);

/*
 * FP64 setup on the device.  An observation of one source: the source's
 * placement (SPHSource, martini/sources/sph_source.py:171-326), the cube's
 * WCS (DataCube, martini/datacube.py:426-562), the kernel tiers
 * (_AdaptiveKernel, martini/sph_kernels.py:1207-1274; one tier for a fixed
 * kernel) and the spectral model (martini/spectral_models.py).  Angles in
 * degrees, distances Mpc, velocities km/s (m/s for the WCS values).
 */
typedef struct mi_observation {
    double rotation[9];      /* source rotation applied to xyz and vxyz first (row-major) */
    double sky_rotation[9];  /* R_z(ra) R_y(-dec) of the source position (row-major) */
    double pointing[3];      /* unit vector towards (ra, dec) */
    double distance_mpc, vpeculiar_kms, hubble_h;
    double cube_ra_deg, cube_dec_deg, px_size_arcsec;
    double crpix[3];         /* 1-based reference pixels of the padded cube */
    double cdelt[3];         /* deg, deg, m/s (VRAD) or Hz (FREQ) */
    double crval2;           /* spectral reference value, m/s or Hz */
    double e0_ms, de_ms;     /* first velocity channel edge and signed width, m/s */
    double dv_max_kms;       /* max |velocity channel width|, km/s (spectral pruning) */
    double sigma_kms;        /* fixed line width, km/s; < 0 selects thermal sqrt(k_B T / m_p) */
    double out_scale;        /* multiplies amp (1/px^2 gives Jy/arcsec^2) */
    double tier_rescale[3], tier_size[3];  /* _rescale and size_in_fwhm per tier */
    double tier_min[3], tier_max[3];       /* tier valid when min <= h * rescale <= max */
    int32_t tier_kind[3], tier_gidx[3];    /* MI_KERNEL_* and Gaussian table index per tier */
    int32_t n_tiers;         /* 1: fixed kernel; 2-3: first valid tier wins */
    int32_t nx, ny, padx, pady, nc;
    int32_t spectrum;        /* MI_SPECTRUM_* */
    int32_t freq_channels;   /* spectral axis in frequency (radio convention) */
    int32_t prune;           /* apply _BaseMartini._prune_particles (martini/martini.py:168-241) */
    int32_t channel_offset;  /* records address channels channel_offset + [0, nc) of a cube */
    int32_t nc_total;        /* ... with nc_total channels (0: nc); several sources then share
                                one insertion, each in its own channel range */
} mi_observation;

/* Flags reported by mi_prepare (counts[2]). */
#define MI_PREP_INVALID_KERNEL 1u /* some kept particle fails kernel validation */
#define MI_PREP_COORD_LIMIT 2u    /* some pixel coordinate beyond 2^21 */

/* Workspace bytes for mi_prepare over n particles. */
int mi_prepare_workspace_size(int64_t n, size_t* bytes);

/*
 * Sky projection, pixel and channel coordinates, smoothing lengths, kernel
 * tier, pruning and record packing of n particles, all in FP64 on the
 * device (replaces Martini.__init__'s setup, martini/martini.py:118-151, and
 * the record packing that mi_insert consumes).  Inputs are device arrays:
 * d_xyz, d_vxyz (n x 3, kpc and km/s, row-major), d_mHI (Msun), d_hsm (kpc),
 * d_T (K; may be null when sigma_kms >= 0), d_edges (nc + 1 velocity channel
 * edges in m/s; needed for the Dirac spectrum, may be null otherwise).
 * Kept particles are written to d_out in input order (d_out holds n records).
 * d_counts (device, 4 x uint64): kept particles, their cell pairs (the
 * n_pairs that mi_insert_workspace_size and mi_insert take), MI_PREP_* flags,
 * and the pixels of the reference's square masks (martini/martini.py:271-274)
 * summed over kept particles (times nc: the particle-voxel updates).
 * Asynchronous.  MI_ERR_UNSUPPORTED for MI_SPECTRUM_TABLE (the caller
 * evaluates table spectra on the host).
 */
int mi_prepare(const double* d_xyz, const double* d_vxyz, const double* d_mHI,
               const double* d_hsm, const double* d_T, const double* d_edges, int64_t n,
               const mi_observation* obs, mi_particle* d_out, uint64_t* d_counts,
               void* d_workspace, size_t workspace_bytes, void* stream);

/*
 * Beam convolution of a cube: every channel map of d_in ([nx][ny][nc],
 * FP32) convolved with the kx x ky beam image d_kernel (row-major, both odd),
 * 'same' output size with zero boundary, written to d_out (not aliasing
 * d_in).  Replaces the per-channel scipy.signal.fftconvolve(..., mode="same")
 * of Martini.convolve_beam (martini/martini.py:863-901); the caller pads the
 * cube by (kx/2, ky/2) first (DataCube.add_pad, martini/datacube.py:669-708)
 * and drops the pad after.  Any beam size: the image is taken in blocks that
 * fit in LDS.  A separable image (the outer product of two vectors, as a
 * circular or axis-aligned Gaussian beam's is) convolves faster as two calls,
 * with the kx x 1 and the 1 x ky factor: one-dimensional images of up to 4096
 * taps take a pass that reads each pixel's channels contiguously and sums in
 * FP64.  Asynchronous.  MI_ERR_LIMIT for dimensions above 32767.
 */
int mi_convolve_beam(const float* d_in, float* d_out, int32_t nx, int32_t ny, int32_t nc,
                     const float* d_kernel, int32_t kx, int32_t ky, void* stream);

/*
 * The FITS data unit of a cube (Martini.write_fits, martini/martini.py:939-1055):
 * the region [padx, padx + nx) x [pady, pady + ny) of the padded cube d_cube
 * ([nx + 2 padx][ny + 2 pady][nc], FP32 when cube_bits is 32, FP64 when 64),
 * written to d_out in FITS axis order (x fastest, then y, then channel: the
 * reference's array transpose, :1049) as big-endian IEEE float32 (bitpix -32)
 * or float64 (bitpix -64, the reference's default for its FP64 cube),
 * nx * ny * nc * |bitpix| / 8 bytes.  d_range (2 doubles) receives the
 * minimum and the maximum (DATAMIN, DATAMAX).  Asynchronous.  MI_ERR_LIMIT
 * for nx or ny above 32767.
 */
int mi_cube_to_fits(const void* d_cube, int32_t cube_bits, int32_t nx, int32_t ny, int32_t nc,
                    int32_t padx, int32_t pady, int32_t bitpix, void* d_out, double* d_range,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MARTINI_MI355X_H */
