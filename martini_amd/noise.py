"""Noise models (martini/noise.py:1-155).

Same classes and contracts as the reference: ``_BaseNoise(seed)`` with
``rng = numpy.random.default_rng(seed)``, ``reset_rng()`` and the abstract
``generate(datacube, beam)``; ``GaussianNoise(rms, seed)``.  ``rms`` is a plain
float in Jy/beam: the noise level wanted after beam convolution.

The noise is drawn on the host with numpy's generator, exactly as the
reference draws it, so a seed gives the same realisation in both; the cube it
is added to then goes through the GPU beam convolution.
"""

from __future__ import annotations

import numpy as np


class _BaseNoise:
    """A noise model: ``generate`` returns noise of the cube array's shape in
    Jy/beam (noise.py:10-88)."""

    def __init__(self, seed: int | None = 0) -> None:
        self.seed = seed
        self.rng = np.random.default_rng(seed=seed)

    def generate(self, datacube, beam) -> np.ndarray:
        raise NotImplementedError

    def reset_rng(self) -> None:
        """Reset the random number generator to its initial state."""
        self.rng = np.random.default_rng(seed=self.seed)


class GaussianNoise(_BaseNoise):
    """Gaussian noise whose rms after convolution with the beam is ``rms``
    Jy/beam (noise.py:91-155)."""

    def __init__(self, rms: float = 1.0, seed: int | None = 0) -> None:
        self.target_rms = float(rms)
        super().__init__(seed=seed)

    def generate(self, datacube, beam) -> np.ndarray:
        """Noise cube (Jy/beam) of the shape of ``datacube._array``.

        Convolution with a Gaussian beam of standard deviations (sig_maj,
        sig_min) pixels lowers the rms of white noise by about
        2 sqrt(pi sig_maj sig_min); the reference's empirical factor 2.19568
        (noise.py:145-150) replaces the 2, and the pre-convolution rms is set
        accordingly."""
        fwhm_to_sigma = 1.0 / (2.0 * np.sqrt(2.0 * np.log(2.0)))
        sig_maj = beam.bmaj * fwhm_to_sigma / datacube.px_size
        sig_min = beam.bmin * fwhm_to_sigma / datacube.px_size
        rms = self.target_rms * 2.19568 * np.sqrt(np.pi * sig_maj * sig_min)
        return self.rng.normal(scale=rms, size=datacube._array.shape)
