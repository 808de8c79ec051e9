"""Spectral-line plugins.

Same classes and contracts as martini.spectral_models
(martini/spectral_models.py:13-587): ``half_width(source)``,
``init_spectra(source, datacube)``, ``spectra``, ``spec_dtype``, ``ncpu``.
The line profile integrated over each channel is evaluated on the GPU by the
same device function the insertion kernels use, from each particle's
channel-space position (see ``_pack``).  ``init_spectra`` is only needed by
callers that want the N x Nc matrix: insertion computes the spectra on the fly
and never materialises it.
"""

from __future__ import annotations

import numpy as np

from . import _lib

K_B = 1.380649e-23  # J/K
M_P = 1.67262192369e-27  # kg


class _BaseSpectrum:
    """Base of the spectral models (spectral_models.py:13-301)."""

    _mi_kind: int = -1

    def __init__(self, ncpu: int | None = None, spec_dtype=np.float64) -> None:
        self.ncpu = ncpu if ncpu is not None else 1
        self.spectra = None
        self.spec_dtype = spec_dtype

    def half_width(self, source):
        raise NotImplementedError

    def _sigma_kms(self, source) -> np.ndarray:
        """Per-particle line width in km/s (0 for a Dirac line)."""
        return np.broadcast_to(np.asarray(self.half_width(source), float), (source.npart,))

    def init_spectra(self, source, datacube) -> None:
        """N x Nc spectra in Jy (spectral_models.py:63-147), computed on the GPU."""
        from ._pack import pack_spectral

        rec, desc = pack_spectral(source, self, datacube)
        self.channel_edges = datacube.velocity_channel_edges
        self.vmids = np.asarray(source.skycoords.radial_velocity, float)
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
