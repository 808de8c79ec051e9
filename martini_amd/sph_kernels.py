"""SPH kernel plugins.

Same classes, names and contracts as martini.sph_kernels
(martini/sph_kernels.py:51-1959): ``kernel(q)``, ``eval_kernel(r, h)``,
``_px_weight(dij, mask)``, ``_kernel_integral(dij, h, mask)``,
``_validate(sm_lengths, noraise, quiet)``, ``_confirm_validation``, and the
attributes ``sm_lengths``, ``sm_ranges``, ``size_in_fwhm``, ``_rescale``,
``kernel_indices``.

The kernel integrals themselves run on the GPU (``mi_kernel_weights`` of the
C ABI): each concrete class carries a device kind code, and the insertion
kernels evaluate the same device functions per (particle, pixel) pair.  The
3D kernel ``kernel(q)`` stays on the host: it only serves FWHM finding and
``eval_kernel``.  Smoothing lengths are in pixels (plain arrays).
"""

from __future__ import annotations

import numpy as np
from scipy.optimize import fsolve
from scipy.special import erf

from . import _lib

ARCSEC_PER_RAD = 180.0 / np.pi * 3600.0
_GAUSS_SIG = 1.0 / (2.0 * np.sqrt(2.0 * np.log(2.0)))  # FWHM = 1


def find_fwhm(f) -> float:
    """FWHM of a kernel peaked at 0 (sph_kernels.py:14-48)."""
    return 2.0 * float(fsolve(lambda q: f(q) - f(np.zeros(1)) / 2, 0.5)[0])


class _BaseSPHKernel:
    """Base of all kernels (sph_kernels.py:51-337)."""

    _mi_kind: int = -1
    size_in_fwhm = None

    def __init__(self) -> None:
        self._rescale = 1.0

    # ----------------------------------------------------- plugin methods
    def _px_weight(self, dij, mask=Ellipsis):
        """Kernel integral over a pixel at offsets ``dij`` (2, N) pixels, using
        the rescaled smoothing lengths (sph_kernels.py:85-119).  Returns pix^-2."""
        if mask is not Ellipsis:
            rescale = self._rescale if np.isscalar(self._rescale) else self._rescale[mask]
            h = np.asarray(self.sm_lengths)[mask] * rescale
        else:
            h = np.asarray(self.sm_lengths) * self._rescale
        return self._kernel_integral(np.asarray(dij, float), h, mask=mask)

    def _confirm_validation(self, noraise: bool = False, quiet: bool = False):
        """Validate the unscaled smoothing lengths (sph_kernels.py:121-138)."""
        return self._validate(self.sm_lengths, noraise=noraise, quiet=quiet)

    def _validate_error(self, msg, sm_lengths, valid, noraise=False, quiet=False):
        """Report and raise (sph_kernels.py:140-190)."""
        if not quiet:
            s = np.asarray(sm_lengths)
            print(f"    ---------{self.__class__.__name__} VALIDATION---------")
            print("    Median smoothing length: ", np.median(s), "px")
            print("    Minimum smoothing length: ", np.min(s), "px")
            print("    Maximum smoothing length: ", np.max(s), "px")
            print("    Smoothing length histogram (np.histogram):", np.histogram(s))
            print("    ", np.sum(np.logical_not(valid)), "/", s.size,
                  "smoothing lengths fail validation.")
        if not noraise:
            raise RuntimeError(msg)

    def eval_kernel(self, r, h):
        """3D kernel at distance r for FWHM h (sph_kernels.py:192-219)."""
        q = np.array(np.asarray(r, float) / np.asarray(h, float) / self._rescale)
        scalar = q.ndim == 0
        W = self.kernel(np.atleast_1d(q)).astype(float)
        W = W / np.power(np.asarray(h, float) * self._rescale, 3)
        return W.item() if scalar else W

    def _apply_mask(self, mask) -> None:
        self.sm_lengths = np.asarray(self.sm_lengths)[mask]
        self.sm_ranges = np.asarray(self.sm_ranges)[mask]

    def _init_sm_lengths(self, source, datacube) -> None:
        """h in pixels: arctan(hsm / d) / px (sph_kernels.py:235-255)."""
        if source.skycoords is None:
            raise RuntimeError("Initialize source skycoords before initializing smoothing lengths.")
        ang = np.arctan(np.asarray(source.hsm_g, float)
                        / (np.asarray(source.skycoords.distance, float) * 1000.0))
        self.sm_lengths = ang * ARCSEC_PER_RAD / datacube.px_size

    def _init_sm_ranges(self) -> None:
        """Mask half-width ceil(h * size_in_fwhm) (sph_kernels.py:257-262)."""
        self.sm_ranges = np.ceil(np.asarray(self.sm_lengths) * self.size_in_fwhm)

    def _kernel_integral(self, dij, h, mask=Ellipsis):
        """Pixel integral of this kernel, evaluated on the GPU."""
        dij = np.asarray(dij, float)
        gt, gn = self._gauss_params()
        w = _lib.kernel_weights(self._mi_kind, dij[0], dij[1], np.asarray(h, float), gt, gn)
        return w.astype(float).reshape(np.shape(dij[0]))

    # -------------------------------------------------------- device side
    def _gauss_params(self):
        return 0.0, 1.0

    def _device_kernels(self):
        """(concrete kernels tuple, per-particle index into it)."""
        n = np.asarray(self.sm_lengths).size
        return (self,), np.zeros(n, dtype=int)

    def _validate(self, sm_lengths, noraise=False, quiet=False):
        raise NotImplementedError

    def kernel(self, q):
        raise NotImplementedError


class _MinSizeKernel(_BaseSPHKernel):
    """Kernels valid above a minimum size (the `_validate` of sph_kernels.py
    :443-478, :687-721, :860-894, :1046-1080, :1566-1600)."""

    min_valid_size: float = 0.0
    _label = ""

    def _validate(self, sm_lengths, noraise=False, quiet=False):
        valid = np.asarray(sm_lengths) >= self.min_valid_size
        if np.logical_not(valid).any():
            self._validate_error(
                f"martini.sph_kernels.{self.__class__.__name__}._validate:\n"
                f"SPH smoothing lengths must be >= {self.min_valid_size:f} px in size for "
                f"{self._label} kernel integral approximation accuracy within 1%.\nThis check "
                "may be disabled by calling martini.martini.Martini.insert_source_in_cube with "
                "'skip_validation=True', but use this with care.\n",
                sm_lengths, valid, noraise=noraise, quiet=quiet)
        return valid


class _FWHMScaledKernel(_MinSizeKernel):
    def __init__(self) -> None:
        super().__init__()
        fwhm = find_fwhm(lambda r: self.eval_kernel(r, 1))
        self.size_in_fwhm = 1.0 / fwhm
        self._rescale = 1.0 / fwhm


class _WendlandC2Kernel(_FWHMScaledKernel):
    """Wendland C2 (sph_kernels.py:340-478)."""

    min_valid_size = 1.51
    _label = "WendlandC2"
    _mi_kind = _lib.KERNEL_WC2

    def kernel(self, q):
        q = np.asarray(q, float)
        return np.where(q < 1, np.power(1 - q, 4) * (4 * q + 1), 0) * (21 / 2 / np.pi)


class _WendlandC6Kernel(_FWHMScaledKernel):
    """Wendland C6 (sph_kernels.py:481-721)."""

    min_valid_size = 1.29
    _label = "WendlandC6"
    _mi_kind = _lib.KERNEL_WC6

    def kernel(self, q):
        q = np.asarray(q, float)
        return np.where(q < 1, np.power(1 - q, 8) * (1 + 8 * q + 25 * q**2 + 32 * q**3), 0) * (
            1365 / 64 / np.pi)


class _CubicSplineKernel(_FWHMScaledKernel):
    """Cubic spline M4 (sph_kernels.py:724-894)."""

    min_valid_size = 1.16
    _label = "CubicSplineKernel"
    _mi_kind = _lib.KERNEL_CUBIC

    def kernel(self, q):
        q = np.asarray(q, float)
        W = np.where(q < 0.5, 1 - 6 * q**2 + 6 * q**3, 2 * np.power(1 - q, 3))
        W = np.where(q > 1, 0.0, W)
        return W * (8 / np.pi)


class _QuarticSplineKernel(_FWHMScaledKernel):
    """Quartic spline M5 (sph_kernels.py:1402-1600)."""

    min_valid_size = 1.2385
    _label = "QuarticSplineKernel"
    _mi_kind = _lib.KERNEL_QUARTIC

    def kernel(self, q):
        q = np.atleast_1d(np.asarray(q, float))
        W = np.where(q < 1, np.power(1 - q, 4), 0.0)
        W = W - np.where(q < 0.6, 5 * np.power(0.6 - q, 4), 0.0)
        W = W + np.where(q < 0.2, 10 * np.power(0.2 - q, 4), 0.0)
        return W * (15625 / 512 / np.pi)


class _GaussianKernel(_MinSizeKernel):
    """Truncated Gaussian (sph_kernels.py:897-1080)."""

    _label = "GaussianKernel"
    _mi_kind = _lib.KERNEL_GAUSS

    def __init__(self, truncate: float = 3.0) -> None:
        self.truncate = float(truncate)
        t = self.truncate
        if t < 2:
            raise RuntimeError("GaussianKernel with truncation <2sigma will cause large errors in "
                               "total mass.")
        self.min_valid_size = (3.7 if t < 3 else 2.3357 if t < 4 else 1.1288 if t < 5
                               else 0.45 if t < 6 else 0.336)
        self.norm = float(erf(t / np.sqrt(2)) - 2 * t / np.sqrt(2 * np.pi) * np.exp(-t * t / 2))
        super().__init__()
        self.size_in_fwhm = t * _GAUSS_SIG

    def kernel(self, q):
        q = np.asarray(q, float)
        sig = _GAUSS_SIG
        return np.where(q < self.truncate * sig,
                        np.power(sig * np.sqrt(2 * np.pi), -3) * np.exp(-np.power(q / sig, 2) / 2),
                        0) / self.norm

    def _gauss_params(self):
        return self.truncate, self.norm


class DiracDeltaKernel(_BaseSPHKernel):
    """Point-like kernel (sph_kernels.py:1083-1204)."""

    max_valid_size = 0.5
    _mi_kind = _lib.KERNEL_DIRAC

    def __init__(self, size_in_fwhm: float = 1.0) -> None:
        super().__init__()
        self.size_in_fwhm = float(size_in_fwhm)
        self._rescale = 1.0

    def kernel(self, q):
        return np.where(np.asarray(q), 0, np.inf)

    def _validate(self, sm_lengths, noraise=False, quiet=False):
        valid = np.asarray(sm_lengths) <= self.max_valid_size
        if np.logical_not(valid).any():
            self._validate_error(
                "martini.sph_kernels.DiracDeltaKernel._validate:\n"
                f"provided smoothing scale (FWHM) must be <= {self.max_valid_size:f} px in size "
                "for DiracDelta kernel to be a reasonable approximation. Call "
                "martini.martini.Martini.insert_source_in_cube with 'skip_validation=True' to "
                "override at the cost of accuracy, but use this with care.",
                sm_lengths, valid, noraise=noraise, quiet=quiet)
        return valid


class _AdaptiveKernel(_BaseSPHKernel):
    """Per-particle choice among kernels in priority order
    (sph_kernels.py:1207-1399)."""

    def __init__(self, kernels) -> None:
        self.kernels = tuple(kernels)
        super().__init__()

    def _init_sm_lengths(self, source, datacube) -> None:
        """Choose the first kernel valid at h * rescale_k (:1241-1274)."""
        super()._init_sm_lengths(source=source, datacube=datacube)
        h = np.asarray(self.sm_lengths)
        idx = -np.ones(h.shape, dtype=int)
        for ik, K in enumerate(self.kernels):
            ok = K._validate(h * K._rescale, noraise=True, quiet=True)
            idx[(idx == -1) & ok] = ik
        self.kernel_indices = idx
        sizes = np.array([K.size_in_fwhm for K in self.kernels], float)
        rescales = np.array([K._rescale for K in self.kernels], float)
        use = np.where(idx < 0, 0, idx)
        self.size_in_fwhm = sizes[use]
        self._rescale = rescales[use]

    def _apply_mask(self, mask) -> None:
        self.size_in_fwhm = np.asarray(self.size_in_fwhm)[mask]
        self._rescale = np.asarray(self._rescale)[mask]
        self.kernel_indices = self.kernel_indices[mask]
        super()._apply_mask(mask)

    def eval_kernel(self, r, h):
        return self.kernels[0].eval_kernel(r, h)

    def kernel(self, q):
        return self.kernels[0].kernel(q)

    def _kernel_integral(self, dij, h, mask=Ellipsis):
        """Dispatch per chosen kernel; index -1 uses kernels[0] (:1330-1363)."""
        dij = np.asarray(dij, float)
        h = np.asarray(h, float)
        idx = np.asarray(self.kernel_indices)[mask]
        use = np.where(idx < 0, 0, idx)
        out = np.zeros(h.shape)
        for ik in np.unique(use):
            m = use == ik
            out[m] = self.kernels[ik]._kernel_integral(dij[:, m], h[m])
        return out

    def _validate(self, sm_lengths, noraise=False, quiet=False):
        valid = self.kernel_indices >= 0
        if np.logical_not(valid).any():
            self._validate_error(
                "martini.sph_kernels._AdaptiveKernel._validate:\n"
                "Some particles have no kernel candidate for which accuracy passes validation.\n"
                "This check may be disabled by calling martini.martini.Martini."
                "insert_source_in_cube with 'skip_validation=True', but use this with care.\n",
                sm_lengths, valid, noraise=noraise, quiet=quiet)
        return valid

    def _device_kernels(self):
        idx = np.asarray(self.kernel_indices)
        return self.kernels, np.where(idx < 0, 0, idx)


class WendlandC2Kernel(_AdaptiveKernel):
    """WendlandC2 with Dirac and Gaussian(t=6) fallbacks (sph_kernels.py:1603-1660)."""

    def __init__(self) -> None:
        super().__init__((_WendlandC2Kernel(), DiracDeltaKernel(), _GaussianKernel(truncate=6.0)))


class WendlandC6Kernel(_AdaptiveKernel):
    """WendlandC6 with fallbacks (sph_kernels.py:1663-1720)."""

    def __init__(self) -> None:
        super().__init__((_WendlandC6Kernel(), DiracDeltaKernel(), _GaussianKernel(truncate=6.0)))


class CubicSplineKernel(_AdaptiveKernel):
    """Cubic spline with fallbacks (sph_kernels.py:1723-1786)."""

    def __init__(self) -> None:
        super().__init__((_CubicSplineKernel(), DiracDeltaKernel(), _GaussianKernel(truncate=6.0)))


class GaussianKernel(_AdaptiveKernel):
    """Gaussian(t) with Dirac and Gaussian(t=6) fallbacks (sph_kernels.py:1789-1863)."""

    def __init__(self, truncate: float = 3.0) -> None:
        super().__init__((_GaussianKernel(truncate=truncate), DiracDeltaKernel(),
                          _GaussianKernel(truncate=6.0)))


class QuarticSplineKernel(_AdaptiveKernel):
    """Quartic spline with fallbacks (sph_kernels.py:1866-1933)."""

    def __init__(self) -> None:
        super().__init__((_QuarticSplineKernel(), DiracDeltaKernel(),
                          _GaussianKernel(truncate=6.0)))


class AdaptiveKernel:
    """Removed interface: raises, as the reference does (sph_kernels.py:1936-1959)."""

    def __init__(self, *args, **kwargs) -> None:
        raise NotImplementedError(
            "Pre-configured adaptive kernels have been implemented in v2.0.3, see the SPH "
            "Kernels page in the documentation for details. You most likely want to use "
            "WendlandC2Kernel(), WendlandC6Kernel(), CubicSplineKernel() or "
            "QuarticSplineKernel() where you previously used AdaptiveKernel(...).")
