"""Beam convolution on the GPU (mi_convolve_beam) against scipy's FFT
convolution, alone and inside Martini.convolve_beam."""

import numpy as np
import pytest
from scipy.signal import fftconvolve

from martini_amd import Martini, _lib, synthetic
from martini_amd.beams import GaussianBeam
from oracle import cube_c
from oracle import martini_oracle as O
from tests.helpers import FLUX_TOL, VOXEL_TOL, oracle_kernel, parity

pytestmark = pytest.mark.gpu


def _kernel(bmaj_px, bmin_px, bpa):
    from martini_amd import DataCube

    b = GaussianBeam(bmaj=bmaj_px, bmin=bmin_px, bpa=bpa, truncate=4)
    b.init_kernel(DataCube(n_px_x=1, n_px_y=1, n_channels=1, px_size=1.0, channel_width=1.0))
    return b.kernel


@pytest.mark.parametrize("shape,beam", [
    ((70, 61, 13), (5.0, 3.0, 30.0)),   # channels not a multiple of 4, elliptical, rotated
    ((48, 48, 8), (14.0, 14.0, 0.0)),   # a 115 px beam image: 17 LDS bands, wider than the map
    ((40, 37, 5), (40.0, 25.0, 35.0)),  # a 323 px image: row blocks of 5 by column blocks of 128
])
def test_convolution_matches_fftconvolve(shape, beam):
    import torch

    rng = np.random.default_rng(7)
    cube = rng.random(shape) * (rng.random(shape) > 0.7)
    K = _kernel(*beam)
    d_in = torch.from_numpy(cube.astype(np.float32)).cuda()
    d_out = torch.empty_like(d_in)
    d_k = torch.from_numpy(K.astype(np.float32)).cuda()
    _lib.convolve_beam(d_in, d_out, *shape, d_k, K.shape[0], K.shape[1])
    gpu = d_out.cpu().numpy().astype(float)
    ref = np.stack([fftconvolve(cube[:, :, c], K, mode="same") for c in range(shape[2])], axis=2)
    err = np.abs(gpu - ref).max() / np.abs(ref).max()
    print(f"{shape} beam {beam} kernel {K.shape}: max err {err:.2e}")
    assert err <= 1e-6


def test_convolve_beam_pipeline_matches_oracle():
    """Insertion into the padded cube, beam convolution, unpadding and
    Jy/beam conversion against the FP64 oracle with scipy's FFT convolution."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    n = 4000
    beam = GaussianBeam(bmaj=40.0, bmin=25.0, bpa=35.0)
    src, dc, k, s = synthetic.build(w, n)
    m = Martini(source=src, datacube=dc, beam=beam, sph_kernel=k, spectral_model=s, quiet=True)
    m.insert_source_in_cube(method="plain")
    m.convolve_beam()
    gpu = m.datacube._array
    padx, pady = beam.needs_pad()
    a = w.arrays(n)
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys, padx=padx, pady=pady)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                  kernel=oracle_kernel(w.kernel), rotation=w.rotation.as_matrix())
    ref = O.convolve_beam(cube_c.cube(P) / w.px_size**2, beam.kernel, padx, pady) * beam.area
    assert gpu.shape == ref.shape == (w.n_px, w.n_px, w.n_channels)
    voxel, flux = parity(gpu, ref)
    print(f"beam pipeline: voxel {voxel:.3e} flux {flux:.3e} (beam image {beam.kernel.shape})")
    assert voxel <= VOXEL_TOL and flux <= FLUX_TOL
    assert m.datacube._array_unit == "Jy/beam" and (m.datacube.padx, m.datacube.pady) == (0, 0)


@pytest.mark.parametrize("beam", [(5.0, 3.0, 0.0), (14.0, 14.0, 0.0)])
def test_separable_passes_match_fftconvolve(beam):
    """An axis-aligned or circular beam image is the outer product of two
    vectors (martini.beam_factors); its two one-dimensional passes, a kx x 1
    and a 1 x ky image through the same entry point, match the 2D FFT
    convolution."""
    import torch

    from martini_amd.martini import beam_factors

    rng = np.random.default_rng(11)
    shape = (57, 64, 9)
    cube = rng.random(shape) * (rng.random(shape) > 0.6)
    K = _kernel(*beam)
    u, v = beam_factors(K)
    d_in = torch.from_numpy(cube.astype(np.float32)).cuda()
    d_tmp, d_out = torch.empty_like(d_in), torch.empty_like(d_in)
    _lib.convolve_beam(d_in, d_tmp, *shape, torch.from_numpy(u.astype(np.float32)).cuda(),
                       u.size, 1)
    _lib.convolve_beam(d_tmp, d_out, *shape, torch.from_numpy(v.astype(np.float32)).cuda(), 1,
                       v.size)
    gpu = d_out.cpu().numpy().astype(float)
    ref = np.stack([fftconvolve(cube[:, :, c], K, mode="same") for c in range(shape[2])], axis=2)
    err = np.abs(gpu - ref).max() / np.abs(ref).max()
    print(f"separable {beam} kernel {K.shape}: max err {err:.2e}")
    assert err <= 1e-6


def test_circular_beam_pipeline_matches_oracle():
    """A circular beam goes through the separable passes, starting from the
    inserted cube still on the device; the result matches the FP64 oracle and
    equals, bit for bit, the same convolution started from the host cube."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    n = 4000
    out = {}
    for resident in (True, False):
        beam = GaussianBeam(bmaj=30.0, bmin=30.0, bpa=0.0)
        src, dc, k, s = synthetic.build(w, n)
        m = Martini(source=src, datacube=dc, beam=beam, sph_kernel=k, spectral_model=s,
                    quiet=True)
        m.insert_source_in_cube(method="mfma")
        assert m._resident is not None
        if not resident:
            m._resident = None  # the host cube is uploaded instead
        m.convolve_beam()
        out[resident] = m.datacube._array
    assert np.array_equal(out[True], out[False])
    padx, pady = beam.needs_pad()
    a = w.arrays(n)
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys, padx=padx, pady=pady)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                  kernel=oracle_kernel(w.kernel), rotation=w.rotation.as_matrix())
    ref = O.convolve_beam(cube_c.cube(P) / w.px_size**2, beam.kernel, padx, pady) * beam.area
    voxel, flux = parity(out[True], ref)
    print(f"circular beam pipeline: voxel {voxel:.3e} flux {flux:.3e} (image {beam.kernel.shape})")
    assert voxel <= VOXEL_TOL and flux <= FLUX_TOL


@pytest.mark.parametrize("shape,k,axis", [
    ((33, 29, 7), 9, 0),      # channels not a multiple of 4: the FP32-lane pass
    ((33, 29, 7), 9, 1),
    ((20, 24, 12), 1, 0),     # a one-tap image: a scaled copy
    ((40, 36, 8), 61, 1),     # an image longer than the map: zero boundary on both ends
    ((9, 4100, 4), 4101, 1),  # longer than the one-dimensional pass holds: the 2D kernel
])
def test_one_dimensional_images(shape, k, axis):
    """kx x 1 and 1 x ky images (the separable passes) against a direct FP64
    sum, including the cases the one-dimensional kernel hands to the 2D one."""
    import torch

    rng = np.random.default_rng(k + axis)
    cube = rng.random(shape)
    taps = rng.random(k)
    taps /= taps.sum()
    d_in = torch.from_numpy(cube.astype(np.float32)).cuda()
    d_out = torch.empty_like(d_in)
    d_k = torch.from_numpy(taps.astype(np.float32)).cuda()
    kx, ky = (k, 1) if axis == 0 else (1, k)
    _lib.convolve_beam(d_in, d_out, *shape, d_k, kx, ky)
    gpu = d_out.cpu().numpy().astype(float)
    c32 = cube.astype(np.float32).astype(float)
    t32 = taps.astype(np.float32).astype(float)
    n = shape[axis]
    ref = np.zeros(shape)
    for a in range(k):  # out[l] += K[a] in[l + k/2 - a], zero outside
        s = k // 2 - a
        lo, hi = max(0, -s), min(n, n - s)
        if lo >= hi:
            continue
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        dst[axis] = slice(lo, hi)
        src[axis] = slice(lo + s, hi + s)
        ref[tuple(dst)] += t32[a] * c32[tuple(src)]
    err = np.abs(gpu - ref).max() / np.abs(ref).max()
    print(f"{shape} k={k} axis={axis}: max err {err:.2e}")
    assert err <= 1e-6


def test_one_dimensional_pass_unaligned_buffers():
    """Cube pointers off 16-byte alignment take the FP32-lane pass and give
    the same result as the aligned vector pass."""
    import torch

    rng = np.random.default_rng(3)
    shape = (24, 20, 16)
    cube = rng.random(shape).astype(np.float32)
    taps = torch.from_numpy(rng.random(7).astype(np.float32)).cuda()
    base_in = torch.zeros(cube.size + 1, dtype=torch.float32, device="cuda")
    base_out = torch.zeros(cube.size + 1, dtype=torch.float32, device="cuda")
    base_in[1:] = torch.from_numpy(cube.ravel()).cuda()
    d_in, d_out = base_in[1:], base_out[1:]  # 4-byte offset
    _lib.convolve_beam(d_in, d_out, *shape, taps, 1, 7)
    aligned_in = torch.from_numpy(cube).cuda()
    aligned_out = torch.empty_like(aligned_in)
    _lib.convolve_beam(aligned_in, aligned_out, *shape, taps, 1, 7)
    assert torch.equal(d_out.reshape(shape), aligned_out)
