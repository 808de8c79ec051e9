"""DataCube: cube geometry, spectral channels and the FP64 host array.

Mirrors martini.datacube.DataCube (martini/datacube.py:49-875) for the parts
the insertion path uses: construction (:157-215), the TAN + VRAD/FREQ WCS
(:426-486), channel edges and mids (:488-562), padding (:669-729) and copy
(:731-759).  astropy is not required: angles are in degrees (ra, dec) or
arcseconds (px_size), velocities in km/s, frequencies in Hz.  Pass
``freq_channels=True`` for a cube evenly spaced in frequency, with
``channel_width`` in Hz (``spectral_centre`` stays a velocity in km/s, and is
converted with the radio convention, as the reference does).
"""

from __future__ import annotations

import numpy as np

HIfreq = 1.420405751e9  # Hz, martini/datacube.py:21
C_LIGHT = 299792458.0  # m/s


class DataCube:
    """Description of the mock cube and its FP64 array (Jy/pix^2 until filled,
    Jy/arcsec^2 after insertion)."""

    def __init__(self, *, n_px_x: int, n_px_y: int, n_channels: int, px_size: float,
                 channel_width: float, spectral_centre: float = 0.0, ra: float = 0.0,
                 dec: float = 0.0, stokes_axis: bool = False, coordinate_frame=None,
                 specsys: str = "icrs", freq_channels: bool = False) -> None:
        if coordinate_frame not in (None, "icrs", "ICRS"):
            raise NotImplementedError("only the ICRS frame is supported")
        if specsys.lower() != "icrs":
            raise NotImplementedError("only specsys='icrs' is supported")
        self.stokes_axis = bool(stokes_axis)
        self.coordinate_frame = "icrs"
        self.specsys = "icrs"
        self.n_px_x, self.n_px_y, self.n_channels = int(n_px_x), int(n_px_y), int(n_channels)
        self.px_size = float(px_size)  # arcsec
        self._freq_channel_mode = bool(freq_channels)
        self.channel_width = abs(float(channel_width))  # km/s, or Hz in frequency mode
        if self._freq_channel_mode:
            self.spectral_centre = HIfreq * (1.0 - float(spectral_centre) * 1e3 / C_LIGHT)  # Hz
        else:
            self.spectral_centre = float(spectral_centre)  # km/s
        self.ra, self.dec = float(ra), float(dec)
        self.padx = 0
        self.pady = 0
        self._array = np.zeros(self._shape())
        self._array_unit = "Jy/pix2"

    def _shape(self):
        shape = (self.n_px_x + 2 * self.padx, self.n_px_y + 2 * self.pady, self.n_channels)
        return shape + (1,) if self.stokes_axis else shape

    # -------------------------------------------------------------- WCS
    @property
    def crpix(self) -> np.ndarray:
        """1-based reference pixel (datacube.py:458-462)."""
        return np.array([self.n_px_x / 2.0 + 0.5 + self.padx, self.n_px_y / 2.0 + 0.5 + self.pady,
                         self.n_channels / 2.0 + 0.5])

    @property
    def cdelt(self) -> np.ndarray:
        """Pixel steps: degrees, degrees, m/s (VRAD, negative) or Hz (FREQ)."""
        spec = self.channel_width if self._freq_channel_mode else -self.channel_width * 1e3
        return np.array([-self.px_size / 3600.0, self.px_size / 3600.0, spec])

    @property
    def crval(self) -> np.ndarray:
        spec = self.spectral_centre if self._freq_channel_mode else self.spectral_centre * 1e3
        return np.array([self.ra, self.dec, spec])

    def _spectral_world(self, pix0):
        return self.crval[2] + self.cdelt[2] * (np.asarray(pix0, float) + 1.0 - self.crpix[2])

    @property
    def channel_edges(self) -> np.ndarray:
        """Channel edges in the cube's spectral unit (m/s or Hz), datacube.py:513-536."""
        return self._spectral_world(np.arange(self.n_channels + 1) - 0.5)

    @property
    def channel_mids(self) -> np.ndarray:
        return self._spectral_world(np.arange(self.n_channels))

    def _to_velocity(self, v):
        return C_LIGHT * (1.0 - v / HIfreq) if self._freq_channel_mode else v

    def _to_frequency(self, v):
        return v if self._freq_channel_mode else HIfreq * (1.0 - v / C_LIGHT)

    @property
    def velocity_channel_edges(self) -> np.ndarray:
        """Edges in m/s (radio convention), datacube.py:551-562."""
        return self._to_velocity(self.channel_edges)

    @property
    def velocity_channel_mids(self) -> np.ndarray:
        return self._to_velocity(self.channel_mids)

    @property
    def frequency_channel_edges(self) -> np.ndarray:
        return self._to_frequency(self.channel_edges)

    @property
    def frequency_channel_mids(self) -> np.ndarray:
        return self._to_frequency(self.channel_mids)

    def world_to_pixel(self, ra_deg, dec_deg, vlos_kms) -> np.ndarray:
        """0-based (x, y, channel) of sky positions: gnomonic (TAN) projection about
        (ra, dec) with RA increasing to the left, then the linear spectral axis
        (the WCS that SPHSource._init_pixcoords uses, sph_source.py:328-362).
        Points on the far hemisphere map to NaN."""
        a0, d0 = np.deg2rad(self.ra), np.deg2rad(self.dec)
        a = np.deg2rad(np.asarray(ra_deg, float))
        d = np.deg2rad(np.asarray(dec_deg, float))
        cos_d, sin_d = np.cos(d), np.sin(d)
        cos_da = np.cos(a - a0)
        den = sin_d * np.sin(d0) + cos_d * np.cos(d0) * cos_da
        front = den > 0
        with np.errstate(invalid="ignore", divide="ignore"):
            xi = np.where(front, np.rad2deg(cos_d * np.sin(a - a0) / den), np.nan)
            eta = np.where(front, np.rad2deg((sin_d * np.cos(d0) - cos_d * np.sin(d0) * cos_da) / den),
                           np.nan)
        crpix, cdelt, crval = self.crpix, self.cdelt, self.crval
        x = crpix[0] - 1.0 + xi / cdelt[0]
        y = crpix[1] - 1.0 + eta / cdelt[1]
        v_ms = np.asarray(vlos_kms, float) * 1e3
        # radio convention; channel values are frequencies in FREQ mode
        spec = HIfreq * (1.0 - v_ms / C_LIGHT) if self._freq_channel_mode else v_ms
        z = crpix[2] - 1.0 + (spec - crval[2]) / cdelt[2]
        return np.vstack((x, y, z))

    def pixel_to_world(self, x, y, z):
        """World coordinates of 0-based pixel positions, the inverse of
        world_to_pixel (WCS.all_pix2world with origin 0, as write_hdf5 uses it,
        martini.py:1229-1244): right ascension in [0, 360) and declination in
        degrees through the inverse gnomonic projection about (ra, dec), and
        the spectral value in the axis unit (m/s, or Hz in frequency mode)."""
        crpix, cdelt, crval = self.crpix, self.cdelt, self.crval
        xi = np.deg2rad((np.asarray(x, float) + 1.0 - crpix[0]) * cdelt[0])
        eta = np.deg2rad((np.asarray(y, float) + 1.0 - crpix[1]) * cdelt[1])
        d0 = np.deg2rad(self.dec)
        den = np.cos(d0) - eta * np.sin(d0)
        ra = np.rad2deg(np.deg2rad(self.ra) + np.arctan2(xi, den)) % 360.0
        dec = np.rad2deg(np.arctan2(np.sin(d0) + eta * np.cos(d0), np.hypot(xi, den)))
        return ra, dec, self._spectral_world(z)

    # ---------------------------------------------------------- padding
    def add_pad(self, pad) -> None:
        """Grow the spatial extent by pad pixels per side (datacube.py:669-708)."""
        if self.padx > 0 or self.pady > 0:
            raise RuntimeError("Tried to add padding to already padded datacube array.")
        old = self._array
        self.padx, self.pady = int(pad[0]), int(pad[1])
        self._array = np.zeros(self._shape())
        xs = slice(self.padx, -self.padx) if self.padx > 0 else slice(None)
        ys = slice(self.pady, -self.pady) if self.pady > 0 else slice(None)
        self._array[xs, ys, ...] = old

    def drop_pad(self) -> None:
        """Remove the padding (datacube.py:710-729)."""
        if self.padx == 0 and self.pady == 0:
            return
        self._array = self._array[self.padx:-self.padx, self.pady:-self.pady, ...]
        self.padx, self.pady = 0, 0

    def copy(self) -> "DataCube":
        """Independent copy (datacube.py:731-759)."""
        c = DataCube.__new__(DataCube)
        c.__dict__.update(self.__dict__)
        c._array = self._array.copy()
        return c

    # -------------------------------------------------------- iteration
    @property
    def spatial_slices(self):
        a = self._array[..., 0] if self.stokes_axis else self._array
        return iter(a.transpose((2, 0, 1)))

    channel_maps = spatial_slices

    @property
    def spectra(self):
        a = self._array[..., 0] if self.stokes_axis else self._array
        return iter(a.reshape(-1, self.n_channels))

    def __repr__(self) -> str:
        return (f"DataCube({self.n_px_x}x{self.n_px_y}x{self.n_channels}, px={self.px_size}\", "
                f"pad=({self.padx},{self.pady}))")
