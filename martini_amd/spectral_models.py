"""Spectral-line plugins.

Same classes and contracts as martini.spectral_models
(martini/spectral_models.py:13-587): ``half_width(source)``,
``init_spectra(source, datacube)``, ``spectra``, ``spec_dtype``, ``ncpu``.
The line profile integrated over each channel is evaluated on the GPU by the
same device function the insertion kernels use, from each particle's
channel-space position (see ``_pack``).  ``init_spectra`` is only needed by
callers that want the N x Nc matrix: insertion computes the spectra on the fly
and never materialises it.

Any other spectral model is a subclass of ``_BaseSpectrum`` that implements
``half_width(source)`` and ``spectral_function(a, b, vmids, extra_data=None)``
(and optionally ``get_spectral_function_extra_data``), as in the reference
(spectral_models.py:203-301).  Velocities are plain floats in km/s: ``a`` and
``b`` the lower and upper channel edges (shape (1, Nc)), ``vmids`` the
particles' line-of-sight velocities (shape (N, 1)); the function returns the
fraction of each particle's line in each channel.  Such a model is evaluated
on the host in FP64, as the reference's init_spectra does, into an N x Nc
table of spectra that the insertion kernels read from HBM (MI_SPECTRUM_TABLE).
"""

from __future__ import annotations

import numpy as np

from . import _lib

K_B = 1.380649e-23  # J/K
M_P = 1.67262192369e-27  # kg


class _BaseSpectrum:
    """Base of the spectral models (spectral_models.py:13-301)."""

    _mi_kind: int = _lib.SPECTRUM_TABLE  # user models: spectra from spectral_function

    def __init__(self, ncpu: int | None = None, spec_dtype=np.float64) -> None:
        self.ncpu = ncpu if ncpu is not None else 1
        self.spectra = None
        self.spec_dtype = spec_dtype

    def half_width(self, source):
        raise NotImplementedError

    def spectral_function(self, a, b, vmids, extra_data=None):
        """Fraction of each line between channel edges a and b (km/s)
        (spectral_models.py:216-251)."""
        raise NotImplementedError

    def get_spectral_function_extra_data(self, source, datacube, mask=np.s_[...]):
        """Extra per-particle data for spectral_function (spectral_models.py:253-301)."""
        return None

    def _sigma_kms(self, source) -> np.ndarray:
        """Per-particle line width in km/s (0 for a Dirac line)."""
        return np.broadcast_to(np.asarray(self.half_width(source), float), (source.npart,))

    def evaluate_spectra(self, source, datacube, mask=np.s_[...]) -> np.ndarray:
        """The fractions spectral_function gives for the channels (N x Nc),
        with the lower and upper edges of each channel taken from the
        monotonic edge sequence (spectral_models.py:149-200)."""
        edges = np.asarray(datacube.velocity_channel_edges, float) / 1e3  # km/s
        steps = np.diff(edges)
        if np.all(steps > 0):
            lower, upper = edges[:-1], edges[1:]
        elif np.all(steps < 0):
            lower, upper = edges[1:], edges[:-1]
        else:
            raise ValueError("Channel edges are not monotonic sequence.")
        vmids = np.asarray(source.skycoords.radial_velocity, float)[mask]
        extra = self.get_spectral_function_extra_data(source, datacube, mask=mask)
        return np.asarray(self.spectral_function(lower[np.newaxis, :].astype(self.spec_dtype),
                                                 upper[np.newaxis, :].astype(self.spec_dtype),
                                                 vmids[:, np.newaxis].astype(self.spec_dtype),
                                                 extra_data=extra), float)

    def _host_spectra(self, source, datacube) -> np.ndarray:
        """N x Nc spectra in Jy, FP64 on the host, as the reference's init_spectra
        (spectral_models.py:63-147): fraction * mHI / d^2 / |dv| / 2.36e5."""
        frac = self.evaluate_spectra(source, datacube)
        widths = np.abs(np.diff(np.asarray(datacube.velocity_channel_edges, float))) / 1e3
        A = np.asarray(source.mHI_g, float) * np.asarray(source.skycoords.distance, float) ** -2.0
        return A[:, np.newaxis] * frac / widths[np.newaxis, :] / 2.36e5

    def init_spectra(self, source, datacube) -> None:
        """N x Nc spectra in Jy (spectral_models.py:63-147): computed on the GPU
        for the Gaussian and Dirac lines, on the host from spectral_function
        for other models."""
        self.channel_edges = datacube.velocity_channel_edges
        self.vmids = np.asarray(source.skycoords.radial_velocity, float)
        if self._mi_kind == _lib.SPECTRUM_TABLE:
            self.spectra = self._host_spectra(source, datacube).astype(self.spec_dtype)
            return
        from ._pack import pack_spectral

        rec, desc = pack_spectral(source, self, datacube)
        self.spectra = _lib.spectra(rec, desc).astype(self.spec_dtype)


class GaussianSpectrum(_BaseSpectrum):
    """Gaussian line of fixed width (km/s) or thermal width sqrt(k_B T / m_p)
    (spectral_models.py:304-485)."""

    _mi_kind = _lib.SPECTRUM_GAUSS

    def __init__(self, sigma=7.0, ncpu: int | None = None, spec_dtype=np.float64) -> None:
        self.sigma_mode = sigma
        super().__init__(ncpu=ncpu, spec_dtype=spec_dtype)

    def half_width(self, source):
        """Velocity dispersion in km/s (spectral_models.py:465-485)."""
        if isinstance(self.sigma_mode, str):
            if self.sigma_mode != "thermal":
                raise ValueError("sigma must be a velocity in km/s or 'thermal'")
            return np.sqrt(K_B * np.asarray(source.T_g, float) / M_P) / 1000.0
        return float(self.sigma_mode)


class DiracDeltaSpectrum(_BaseSpectrum):
    """Zero-width line (spectral_models.py:488-587)."""

    _mi_kind = _lib.SPECTRUM_DIRAC

    def half_width(self, source):
        return 0.0
