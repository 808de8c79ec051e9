"""Noise (martini/noise.py, Martini.add_noise martini/martini.py:903-937),
following the reference's tests/test_noise.py: shape, seeded repeatability,
unseeded variety, reset_rng, and the post-convolution rms (on the GPU).
The realisation is numpy's default_rng(seed).normal, as in the reference."""

import warnings

import numpy as np
import pytest

from martini_amd import DataCube, Martini, SPHSource
from martini_amd.beams import GaussianBeam
from martini_amd.noise import GaussianNoise, _BaseNoise
from martini_amd.sph_kernels import CubicSplineKernel
from martini_amd.spectral_models import GaussianSpectrum


def _cube():
    return DataCube(n_px_x=256, n_px_y=256, n_channels=64, px_size=15.0, channel_width=4.0)


def test_noise_shape():
    dc = _cube()
    assert GaussianNoise(rms=1.0).generate(dc, GaussianBeam()).shape == dc._array.shape


def test_noise_is_numpys_normal_with_the_reference_rms():
    """rms before convolution = target * 2.19568 * sqrt(pi sig_maj sig_min)
    (noise.py:145-150), sig = FWHM / (2 sqrt(2 ln 2)) in pixels."""
    dc = DataCube(n_px_x=16, n_px_y=12, n_channels=8, px_size=15.0, channel_width=4.0)
    beam = GaussianBeam(bmaj=20.0, bmin=15.0)
    got = GaussianNoise(rms=3e-5, seed=11).generate(dc, beam)
    s = 2.0 * np.sqrt(2.0 * np.log(2.0))
    rms = 3e-5 * 2.19568 * np.sqrt(np.pi * (20.0 / s / 15.0) * (15.0 / s / 15.0))
    np.testing.assert_array_equal(got, np.random.default_rng(11).normal(scale=rms,
                                                                        size=(16, 12, 8)))


def test_noise_seed_and_reset():
    dc, beam = _cube(), GaussianBeam()
    a = GaussianNoise(rms=1.0, seed=0).generate(dc, beam)
    b = GaussianNoise(rms=1.0, seed=0).generate(dc, beam)
    np.testing.assert_array_equal(a, b)
    assert not np.allclose(GaussianNoise(rms=1.0, seed=None).generate(dc, beam),
                           GaussianNoise(rms=1.0, seed=None).generate(dc, beam))
    g = GaussianNoise(rms=1.0, seed=0)
    first = g.generate(dc, beam)
    assert not np.allclose(first, g.generate(dc, beam))
    g.reset_rng()
    np.testing.assert_array_equal(first, g.generate(dc, beam))
    with pytest.raises(NotImplementedError):
        _BaseNoise().generate(dc, beam)


def _m_init(noise=True, beam=True):
    """The reference's m_init fixture (tests/conftest.py:470-510), in plain units."""
    rng = np.random.default_rng(1)
    n = 50
    src = SPHSource(distance=3.0, mHI_g=np.full(n, 1e4), xyz_g=rng.normal(0.0, 0.2, (n, 3)),
                    vxyz_g=rng.normal(0.0, 5.0, (n, 3)), hsm_g=np.full(n, 1.0),
                    T_g=np.full(n, 1e4))
    dc = DataCube(n_px_x=16, n_px_y=16, n_channels=16, px_size=15.0, channel_width=4.0,
                  spectral_centre=210.0)
    return Martini(source=src, datacube=dc,
                   beam=GaussianBeam(bmaj=20.0, bmin=15.0) if beam else None,
                   noise=GaussianNoise(rms=1.0e-9, seed=0) if noise else None,
                   sph_kernel=CubicSplineKernel(), spectral_model=GaussianSpectrum(),
                   quiet=True)


def test_add_noise_units_and_skips():
    m = _m_init()
    before = m.datacube._array.copy()
    m.add_noise()  # not inserted yet: Jy/beam -> Jy/arcsec^2 -> Jy/pix^2
    m.noise.reset_rng()
    raw = m.noise.generate(m.datacube, m.beam)
    np.testing.assert_allclose(m.datacube._array - before,
                               raw / m.beam.area * m.datacube.px_size**2, rtol=1e-15)
    for kw in ({"noise": False}, {"beam": False}):
        m2 = _m_init(**kw)
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            m2.add_noise()
        assert any("Skipping noise" in str(x.message) for x in w)
        assert not m2.datacube._array.any()


@pytest.mark.gpu
def test_noise_amplitude_after_convolution():
    """The reference's test_noise_amplitude: insert, zero the cube, add noise,
    convolve with the beam (on the GPU): rms within 10% of the target."""
    m = _m_init()
    m.insert_source_in_cube()
    m.datacube._array[...] = 0.0
    m.add_noise()
    m.convolve_beam()
    measured = np.sqrt(np.mean(m.datacube._array**2))
    print(f"noise rms {measured:.4e} target {m.noise.target_rms:.4e}")
    assert measured == pytest.approx(m.noise.target_rms, rel=0.1)
