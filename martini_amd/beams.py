"""Telescope beams (martini/beams.py): the beam image and its padding.

Same classes and contracts as the reference: ``_BaseBeam`` with
``f_kernel``, ``kernel_size_px``, ``init_beam_header``, ``init_kernel``,
``needs_pad`` and ``area``; ``GaussianBeam(bmaj, bmin, bpa, truncate)``.
Angles are plain floats: bmaj and bmin in arcsec, bpa in degrees.

The beam image is built on the host in FP64 exactly as the reference builds
it (martini/beams.py:103-159): the beam function sampled every 0.1 pixel, a
bicubic spline through the samples, and the spline's integral over each
pixel.  ``Martini.convolve_beam`` convolves the cube with it on the GPU
(``mi_convolve_beam``).
"""

from __future__ import annotations

import numpy as np
import scipy.interpolate


class _BaseBeam:
    """A radio telescope beam (martini/beams.py:11-194)."""

    def __init__(self, bmaj: float | None = 15.0, bmin: float | None = None,
                 bpa: float | None = 0.0) -> None:
        self.bmaj = bmaj
        self.bmin = bmin if bmin is not None else bmaj
        self.bpa = bpa
        self.px_size = None
        self.kernel = None
        # FWHM -> Gaussian-equivalent area: pi bmaj bmin / (4 ln 2) (arcsec^2)
        self.area = (np.pi * self.bmaj * self.bmin) / 4 / np.log(2) if bmaj is not None else None

    def needs_pad(self) -> tuple[int, int]:
        """Padding of the cube (x, y) the beam needs against edge effects."""
        if self.kernel is None:
            raise RuntimeError("Beam kernel not initialized.")
        return self.kernel.shape[0] // 2, self.kernel.shape[1] // 2

    def init_kernel(self, datacube) -> None:
        """The beam image: integral of the beam over each pixel (beams.py:103-159)."""
        self.px_size = datacube.px_size
        self.vel = datacube.spectral_centre
        self.ra = datacube.ra
        self.dec = datacube.dec
        if self.bmaj is None or self.bmin is None or self.bpa is None:
            self.init_beam_header()
            self.area = (np.pi * self.bmaj * self.bmin) / 4 / np.log(2)
        npx_x, npx_y = self.kernel_size_px()
        px = float(self.px_size)
        px_edges_x = np.arange(-npx_x - 0.5, npx_x + 0.50001, 1) * px
        px_edges_y = np.arange(-npx_y - 0.5, npx_y + 0.50001, 1) * px
        fine_x = np.arange(-npx_x - 0.5, npx_x + 0.501, 0.1) * px
        fine_y = np.arange(-npx_y - 0.5, npx_y + 0.501, 0.1) * px
        rbs = scipy.interpolate.RectBivariateSpline(
            fine_x, fine_y, self.f_kernel()(*np.meshgrid(fine_x, fine_y, indexing="ij")),
            kx=3, ky=3)
        xgrid, ygrid = np.meshgrid(px_edges_x, px_edges_y)
        self.kernel = np.vectorize(rbs.integral)(xgrid[1:, :-1], xgrid[1:, 1:], ygrid[:-1, 1:],
                                                 ygrid[1:, 1:])

    def f_kernel(self):
        raise NotImplementedError

    def kernel_size_px(self) -> tuple[int, int]:
        raise NotImplementedError

    def init_beam_header(self) -> None:
        raise NotImplementedError


class GaussianBeam(_BaseBeam):
    """Elliptical Gaussian beam (martini/beams.py:197-296): bmaj, bmin FWHM in
    arcsec, bpa east of north in degrees, truncated at ``truncate`` FWHM."""

    def __init__(self, bmaj: float = 15.0, bmin: float | None = None, bpa: float = 0.0,
                 truncate: float = 4.0) -> None:
        self.truncate = truncate
        super().__init__(bmaj=bmaj, bmin=bmin, bpa=bpa)

    def f_kernel(self):
        """Beam amplitude (arcsec^-2) at (RA, Dec) offsets in arcsec."""
        def fwhm_to_sigma(fwhm):
            return fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))

        smaj = fwhm_to_sigma(self.bmaj)
        smin = fwhm_to_sigma(self.bmin)
        pa = np.deg2rad(self.bpa)
        a = np.cos(pa) ** 2 / (2.0 * smin**2) + np.sin(pa) ** 2 / (2.0 * smaj**2)
        b = -np.sin(2.0 * pa) / (4 * smin**2) + np.sin(2.0 * pa) / (4 * smaj**2)
        c = np.sin(pa) ** 2 / (2.0 * smin**2) + np.cos(pa) ** 2 / (2.0 * smaj**2)
        A = 1.0 / (2.0 * np.pi * smin * smaj)
        return lambda x, y: A * np.exp(-a * x**2 - 2.0 * b * x * y - c * y**2)

    def kernel_size_px(self) -> tuple[int, int]:
        """Half-size of the beam image: ceil(bmaj * truncate / px + 1)."""
        size = int(np.ceil(self.bmaj * self.truncate / float(self.px_size) + 1))
        return size, size

    def init_beam_header(self) -> None:
        """Nothing to do: the header is set at construction."""
