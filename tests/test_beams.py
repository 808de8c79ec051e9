"""Beam image (host, FP64) against the reference's known answers
(tests/test_beams.py:13-86) and the closed form of a circular beam."""

import numpy as np
import pytest
from scipy.special import erf

from martini_amd import DataCube, Martini, synthetic
from martini_amd.beams import GaussianBeam
from martini_amd.sph_kernels import CubicSplineKernel
from martini_amd.spectral_models import GaussianSpectrum


def _cube(px=1.0):
    return DataCube(n_px_x=1, n_px_y=1, n_channels=1, px_size=px, channel_width=1.0)


def test_gaussian_beam_size():
    b = GaussianBeam(bmaj=20.0, bmin=10.0, bpa=0.0, truncate=4)
    b.init_kernel(_cube())
    half = int(np.ceil(20.0 * 4 / 1.0)) + 1
    assert b.kernel.shape == (2 * half + 1, 2 * half + 1)
    assert b.needs_pad() == (half, half)


@pytest.mark.parametrize("bpa", [0.0, 90.0])
def test_gaussian_beam_rotation(bpa):
    b = GaussianBeam(bmaj=20.0, bmin=10.0, bpa=bpa, truncate=4)
    b.init_kernel(_cube())
    mid = b.kernel.shape[0] // 2
    if bpa == 0.0:
        assert np.isclose(b.kernel[mid, mid + 5], b.kernel[mid + 10, mid], rtol=1e-2)
    else:
        assert np.isclose(b.kernel[mid, mid + 10], b.kernel[mid + 5, mid], rtol=1e-2)


def test_gaussian_beam_integral():
    b = GaussianBeam(bmaj=10.0, bmin=10.0, bpa=0.0, truncate=4)
    b.init_kernel(_cube())
    assert np.isclose(np.sum(b.kernel), 1.0)


def test_circular_beam_matches_separable_pixel_integral():
    """A circular Gaussian's pixel integral is a product of erf differences;
    the spline integral reproduces it to 1e-4 of the peak."""
    px, fwhm = 2.0, 9.0
    b = GaussianBeam(bmaj=fwhm, bmin=fwhm, bpa=0.0, truncate=4)
    b.init_kernel(_cube(px))
    n = b.kernel.shape[0] // 2
    s = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    e = (np.arange(-n - 0.5, n + 1.0, 1.0)) * px / (np.sqrt(2.0) * s)
    p1 = 0.5 * np.diff(erf(e))
    exact = np.outer(p1, p1)
    assert np.abs(b.kernel - exact).max() <= 1e-4 * exact.max()


def test_area():
    b = GaussianBeam(bmaj=12.0, bmin=6.0)
    assert np.isclose(b.area, np.pi * 12.0 * 6.0 / 4 / np.log(2))


def test_martini_pads_for_the_beam_and_reset_keeps_it():
    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, s = synthetic.build(w, 500)
    beam = GaussianBeam(bmaj=40.0, bmin=30.0, bpa=20.0)
    m = Martini(source=src, datacube=dc, beam=beam, sph_kernel=k, spectral_model=s, quiet=True)
    half = beam.needs_pad()
    assert (m.datacube.padx, m.datacube.pady) == half
    assert m.datacube._array.shape[:2] == (64 + 2 * half[0], 64 + 2 * half[1])
    m.reset()
    assert (m.datacube.padx, m.datacube.pady) == half
    m.datacube.drop_pad()
    m.reset()
    assert (m.datacube.padx, m.datacube.pady) == half
