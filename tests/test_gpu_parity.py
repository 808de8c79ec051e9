"""GPU parity: the HIP insertion path against the FP64 oracle.

Bar: per voxel |gpu - oracle| <= 1e-5 of the oracle cube's peak, and total
flux within 1e-6, at workload sizes the oracle finishes in seconds.
"""

import numpy as np
import pytest

from martini_amd import synthetic
from oracle import cube_c
from oracle import martini_oracle as O
from tests.helpers import FLUX_TOL, VOXEL_TOL, gpu_cube, oracle_cube, oracle_kernel, parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_disc1e4_64px(method):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    ref, P = oracle_cube(w)
    gpu, m = gpu_cube(w, method=method)
    voxel, flux = parity(gpu, ref)
    print(f"disc1e4_64px {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL
    assert flux <= FLUX_TOL


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_disc1e6_256px(method):
    w = synthetic.WORKLOADS["disc1e6_256px"]
    ref, P = oracle_cube(w)
    gpu, m = gpu_cube(w, method=method)
    voxel, flux = parity(gpu, ref)
    print(f"disc1e6_256px {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL
    assert flux <= FLUX_TOL


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_deterministic(method):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    a, _ = gpu_cube(w, method=method)
    b, _ = gpu_cube(w, method=method)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_hetero_smoothing_subsample(method):
    """The heterogeneous-smoothing configuration (h 0.05-50 px, T 1e2 / 1e4 K,
    1024^2 x 64) on a 1e5-particle subsample: tiny Dirac-fallback particles
    and footprints spanning ~100 tiles in one launch."""
    w = synthetic.WORKLOADS["hetero3e6_1024px"]
    n = 100_000
    ref, P = oracle_cube(w, n=n)
    gpu, m = gpu_cube(w, n=n, method=method)
    voxel, flux = parity(gpu, ref)
    print(f"hetero n={n} {method}: voxel {voxel:.3e} flux {flux:.3e} "
          f"r_max {int(P.sm_ranges.max())}")
    assert voxel <= VOXEL_TOL
    assert flux <= FLUX_TOL


def _kernel_cases():
    from martini_amd.sph_kernels import (CubicSplineKernel, DiracDeltaKernel, GaussianKernel,
                                         QuarticSplineKernel, WendlandC6Kernel)
    from martini_amd.spectral_models import DiracDeltaSpectrum, GaussianSpectrum

    return {
        "wc6": (WendlandC6Kernel, GaussianSpectrum, oracle_kernel("wc6"), "gauss", False),
        "quartic": (QuarticSplineKernel, GaussianSpectrum, oracle_kernel("quartic"), "gauss", False),
        "gauss_t3": (lambda: GaussianKernel(truncate=3.0), GaussianSpectrum,
                     oracle_kernel("gauss", truncate=3.0), "gauss", False),
        "dirac_kernel": (DiracDeltaKernel, GaussianSpectrum, O.KernelSpec("dirac"), "gauss", True),
        "dirac_spectrum": (CubicSplineKernel, DiracDeltaSpectrum, oracle_kernel("cubic"), "dirac",
                           False),
    }


@pytest.mark.parametrize("case", ["wc6", "quartic", "gauss_t3", "dirac_kernel", "dirac_spectrum"])
@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_kernel_and_spectrum_kinds(case, method):
    """Every kernel and spectrum kind through the full path, on the 64 px
    disc with 4000 particles."""
    from martini_amd import Martini

    make_k, make_s, oracle_k, oracle_s, skip = _kernel_cases()[case]
    w = synthetic.WORKLOADS["disc1e4_64px"]
    n = 4000
    ref, P = oracle_cube(w, n=n, kernel=oracle_k, spectrum=oracle_s)
    spec = make_s(sigma="thermal") if oracle_s == "gauss" else make_s()
    src, dc, k, s = synthetic.build(w, n, kernel=make_k(), spectrum=spec)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    m.insert_source_in_cube(skip_validation=skip, method=method)
    voxel, flux = parity(m.datacube._array, ref)
    print(f"{case} {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL
    assert flux <= FLUX_TOL


@pytest.mark.parametrize("spectrum", ["gauss", "dirac"])
@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_global_profile(spectrum, method):
    """GlobalProfile's spatially integrated spectrum against the oracle."""
    from martini_amd import GlobalProfile, SPHSource
    from martini_amd.spectral_models import DiracDeltaSpectrum, GaussianSpectrum

    w = synthetic.WORKLOADS["disc1e4_64px"]
    a = w.arrays(4000)
    src = SPHSource(distance=w.distance, rotation=w.rotation, h=0.7, T_g=a.T, mHI_g=a.mHI,
                    xyz_g=a.xyz, vxyz_g=a.vxyz, hsm_g=a.hsm)
    model = GaussianSpectrum(sigma="thermal") if spectrum == "gauss" else DiracDeltaSpectrum()
    gp = GlobalProfile(source=src, spectral_model=model, n_channels=48, channel_width=6.0,
                       spectral_centre=w.vsys, quiet=True)
    gp.insert_source_in_spectrum(method=method)
    ref, kept = O.global_profile(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance,
                                 n_channels=48, channel_width=6.0, spectral_centre=w.vsys,
                                 spectrum=spectrum, rotation=w.rotation.as_matrix())
    err = np.abs(gp.spectrum - ref).max() / np.abs(ref).max()
    flux = abs(gp.spectrum.sum() - ref.sum()) / ref.sum()
    print(f"global profile {spectrum} {method}: max {err:.3e} flux {flux:.3e}, kept {kept}")
    assert gp.source.npart == kept
    assert err <= VOXEL_TOL and flux <= FLUX_TOL


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_frequency_channels(method):
    """A cube with channels evenly spaced in frequency (FREQ axis, radio
    convention): the disc1e4 source, 32 channels of 19 kHz (about 4 km/s)."""
    from martini_amd import DataCube, Martini, SPHSource
    from martini_amd.sph_kernels import CubicSplineKernel
    from martini_amd.spectral_models import GaussianSpectrum

    w = synthetic.WORKLOADS["disc1e4_64px"]
    a = w.arrays()
    df = 19_000.0
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, df, spectral_centre=w.vsys,
                      freq_mode=True)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                  kernel=oracle_kernel("cubic"), rotation=w.rotation.as_matrix())
    ref = cube_c.cube(P) / w.px_size**2
    src = SPHSource(distance=w.distance, rotation=w.rotation, T_g=a.T, mHI_g=a.mHI, xyz_g=a.xyz,
                    vxyz_g=a.vxyz, hsm_g=a.hsm)
    dc = DataCube(n_px_x=w.n_px, n_px_y=w.n_px, n_channels=w.n_channels, px_size=w.px_size,
                  channel_width=df, spectral_centre=w.vsys, freq_channels=True)
    m = Martini(source=src, datacube=dc, sph_kernel=CubicSplineKernel(),
                spectral_model=GaussianSpectrum(sigma="thermal"), quiet=True)
    m.insert_source_in_cube(method=method)
    voxel, flux = parity(m.datacube._array, ref)
    print(f"freq channels {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL
    assert flux <= FLUX_TOL


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_insert_twice_accumulates(method):
    """insert_source_in_cube adds to the cube (martini/martini.py:338, `+=`):
    a second insertion of the same source doubles it exactly (the same FP32
    contribution added twice in FP64), and the Jy/pix^2 -> Jy/arcsec^2
    conversion happens once."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, s = synthetic.build(w, 3000)
    from martini_amd import Martini

    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    m.insert_source_in_cube(method=method)
    once = m.datacube._array.copy()
    assert m.datacube._array_unit == "Jy/arcsec2"
    m.insert_source_in_cube(method=method)
    assert m.datacube._array_unit == "Jy/arcsec2"
    assert once.max() > 0
    assert np.array_equal(m.datacube._array, 2.0 * once)
    assert m._resident is None  # the device cube no longer mirrors the host cube
