"""Martini: the mock-observation orchestrator, insertion path only.

Mirrors martini.martini.Martini / _BaseMartini (martini/martini.py:47-425,
:627-861): the constructor runs the same FP64 setup sequence (sky and pixel
coordinates, smoothing lengths, ranges, pruning; :118-151, :168-241), and
``insert_source_in_cube`` keeps its signature and effect (:827-861): the
cube array is accumulated (+=) and converted to Jy/arcsec^2, and a flux and
mass summary is printed unless quiet.

The pixel loop of the reference (:243-283, :327-362) is replaced by the GPU
path: records packed on the host (``_pack``), then binning, a tile gather
(``method='plain'`` FP32 or ``'mfma'`` BF16x3) and a partial-tile reduction
on the device.  ``ncpu`` and ``progressbar`` are accepted for compatibility
and have no effect.  A beam pads the cube at construction and
``convolve_beam`` convolves it on the GPU (martini.py:138-140, :863-901).
``add_noise`` adds a noise model's realisation (martini.py:903-937), drawn
with numpy's generator as the reference draws it.  ``write_fits`` /
``write_beam_fits`` write FITS files (martini.py:939-1152), the cube's data
unit made on the GPU; ``write_hdf5`` writes the HDF5 file of
martini.py:1154-1366 through the HDF5 C library (``hdf5.py``).
"""

from __future__ import annotations

import warnings

import numpy as np

from . import __version__, _lib, _pack, fits, hdf5
from .insertion import default_inserter

# str(unit) of the cube array, as astropy prints the units the reference uses
_UNIT_STR = {"Jy/beam": "Jy / beam", "Jy/arcsec2": "Jy / arcsec2", "Jy/pix2": "Jy / pix2"}


class _BaseMartini:
    def __init__(self, *, source, datacube, beam=None, noise=None, sph_kernel, spectral_model,
                 quiet: bool = False, _prune_kwargs: dict | None = None) -> None:
        self.quiet = quiet
        self.source = source
        self._datacube = datacube
        self.beam = beam
        self.noise = noise
        self.sph_kernel = sph_kernel
        self.spectral_model = spectral_model
        if self.beam is not None:
            self.beam.init_kernel(self._datacube)
            self._datacube.add_pad(self.beam.needs_pad())
        self.source._init_skycoords()
        self.source._init_pixcoords(self._datacube)
        self.sph_kernel._init_sm_lengths(source=self.source, datacube=self._datacube)
        self.sph_kernel._init_sm_ranges()
        self._prune_particles(**(_prune_kwargs or {}))

    @property
    def datacube(self):
        return self._datacube

    def init_spectra(self) -> None:
        """N x Nc spectra, for callers that want them (martini.py:153-166)."""
        if not self.quiet:
            print("Initializing spectra...")
        self.spectral_model.init_spectra(self.source, self._datacube)
        if not self.quiet:
            print("Spectra initialized.")

    def _prune_particles(self, spatial: bool = True, spectral: bool = True, mass: bool = True,
                         obj_type_str: str = "data cube") -> None:
        """Drop particles that cannot contribute (martini.py:168-241)."""
        src, dc = self.source, self._datacube
        if not self.quiet:
            print(f"Source module contained {src.npart} particles with total HI mass of "
                  f"{src.mHI_g.sum():.2e} solMass.")
        x, y, z = src.pixcoords
        r = np.asarray(self.sph_kernel.sm_ranges, float)
        dv_max = np.max(np.abs(np.diff(dc.velocity_channel_edges))) / 1e3  # km/s
        hw = np.asarray(self.spectral_model.half_width(src), float) / dv_max
        reject = np.zeros(src.npart, bool)
        with np.errstate(invalid="ignore"):
            if spatial:
                reject |= (x + r < 0) | (y + r < 0)
                reject |= x - r > dc.n_px_x + dc.padx * 2
                reject |= y - r > dc.n_px_y + dc.pady * 2
                reject |= np.isnan(x) | np.isnan(y)
            if spectral:
                reject |= z + 4 * hw < 0
                reject |= z - 4 * hw > dc.n_channels
        if mass:
            reject |= src.mHI_g == 0
        keep = ~reject
        src.apply_mask(keep)
        self.sph_kernel._apply_mask(keep)
        if not self.quiet:
            print(f"Pruned particles that will not contribute to {obj_type_str}, {src.npart} "
                  f"particles remaining with total HI mass of {src.mHI_g.sum():.2e} solMass.")

    def _insert_source_in_cube(self, skip_validation: bool = False, progressbar=None,
                               ncpu: int = 1, quiet=None, method: str = "plain") -> None:
        """Fill the cube on the GPU (replaces martini.py:285-407)."""
        self.sph_kernel._confirm_validation(noraise=skip_validation, quiet=self.quiet)
        dc = self._datacube
        px2 = dc.px_size ** 2
        rec, desc = _pack.pack_particles(self.source, self.sph_kernel, self.spectral_model, dc,
                                         out_scale=1.0 / px2)
        # the cube is empty unless something was inserted or added before
        was_empty = not np.any(dc._array)
        d_cube = default_inserter().insert(rec, desc, method=method)
        added = d_cube.cpu().numpy().astype(np.float64).reshape(dc._array.shape[:3])
        if dc._array_unit == "Jy/pix2":
            dc._array = dc._array / px2
            dc._array_unit = "Jy/arcsec2"
        if dc.stokes_axis:
            dc._array[..., 0] += added
        else:
            dc._array += added
        # The device cube equals the host cube while this array object holds
        # exactly the insertion (FP32 values widened to FP64); convolve_beam
        # then starts from it.  The methods of this class replace the array
        # object whenever they change the cube.
        self._resident = (d_cube, dc._array) if was_empty else None
        self._report(quiet)

    def _report(self, quiet) -> None:
        if not ((quiet is None and not self.quiet) or (quiet is not None and not quiet)):
            return
        dc = self._datacube
        pad = (np.s_[dc.padx:-dc.padx, dc.pady:-dc.pady, ...] if dc.padx > 0 and dc.pady > 0
               else np.s_[...])
        flux = np.sum(dc._array[pad] * dc.px_size ** 2)
        dv = np.abs(np.diff(dc.velocity_channel_edges)) / 1e3
        per_chan = (dc._array[pad] * dc.px_size ** 2).sum((0, 1)).squeeze()
        mass = 2.36e5 * self.source.distance ** 2 * np.sum(per_chan * dv)
        nz = dc._array[dc._array > 0]
        print("Source inserted.",
              f"  Flux density in cube: {flux:.2e} Jy",
              f"  Mass in cube (assuming distance {self.source.distance:.2f} Mpc and a spatially "
              f"resolved source): {mass:.2e} solMass",
              f"    [{mass / self.source.input_mass * 100:.0f}% of initial source mass]",
              f"  Maximum pixel: {dc._array.max():.2e} Jy / arcsec2",
              f"  Median non-zero pixel: {np.median(nz) if nz.size else 0.0:.2e} Jy / arcsec2",
              sep="\n")

    def reset(self) -> None:
        """A zero cube of the same geometry, padded for the beam (martini.py:409-425)."""
        self._resident = None
        dc = self._datacube
        dc.padx = dc.pady = 0
        dc._array = np.zeros(dc._shape())
        dc._array_unit = "Jy/pix2"
        if self.beam is not None:
            dc.add_pad(self.beam.needs_pad())


# A beam image is convolved as two one-dimensional passes when its second
# singular value is below this fraction of the first: the outer product of the
# two factors then equals the image to FP64 rounding.
SEPARABLE_RTOL = 1e-12


def beam_factors(kernel):
    """(column factor u, row factor v) with kernel = outer(u, v) when the
    beam image is separable to FP64 precision (a circular or axis-aligned
    Gaussian beam's image is: the reference builds it from a tensor-product
    spline of a separable function), else None."""
    K = np.asarray(kernel, float)
    U, S, Vt = np.linalg.svd(K)
    if S[0] <= 0 or (S.size > 1 and S[1] > SEPARABLE_RTOL * S[0]):
        return None
    u, v = U[:, 0] * np.sqrt(S[0]), Vt[0] * np.sqrt(S[0])
    if u.sum() < 0:
        u, v = -u, -v
    return u, v


class Martini(_BaseMartini):
    """Creates synthetic HI data cubes from SPH particles (martini.py:627-861)."""

    def __init__(self, *, source, datacube, beam=None, noise=None, sph_kernel, spectral_model,
                 quiet: bool = False) -> None:
        super().__init__(source=source, datacube=datacube, beam=beam, noise=noise,
                         sph_kernel=sph_kernel, spectral_model=spectral_model, quiet=quiet)

    def convolve_beam(self) -> None:
        """Convolve every channel map with the beam image, drop the padding and
        convert Jy/arcsec^2 to Jy/beam (martini.py:863-901); on the GPU."""
        from .insertion import default_inserter

        if self.beam is None:
            warnings.warn("Skipping beam convolution, no beam object provided to Martini.")
            return
        dc = self._datacube
        pad = self.beam.needs_pad()
        if dc.padx < pad[0] or dc.pady < pad[1]:
            raise ValueError("datacube padding insufficient for beam convolution (perhaps you "
                             "loaded a datacube state with datacube.load_state that was "
                             "previously initialized by martini with a smaller beam?)")
        ins = default_inserter()
        torch = ins.torch
        arr = dc._array[..., 0] if dc.stokes_axis else dc._array
        nx, ny, nc = arr.shape
        resident = getattr(self, "_resident", None)
        if resident is not None and resident[1] is dc._array:
            d_in = resident[0].view(nx, ny, nc)  # the inserted cube, still on the device
        else:
            d_in = torch.from_numpy(np.ascontiguousarray(arr, np.float32)).to(ins.device)
        self._resident = None
        d_out = torch.empty_like(d_in)
        factors = beam_factors(self.beam.kernel)
        if factors is not None:  # two one-dimensional passes
            d_tmp = torch.empty_like(d_in)
            for d_src, d_dst, f, (kx, ky) in ((d_in, d_tmp, factors[0], (factors[0].size, 1)),
                                              (d_tmp, d_out, factors[1], (1, factors[1].size))):
                d_k = torch.from_numpy(np.ascontiguousarray(f, np.float32)).to(ins.device)
                _lib.convolve_beam(d_src, d_dst, nx, ny, nc, d_k, kx, ky)
        else:
            K = np.ascontiguousarray(self.beam.kernel, np.float32)
            d_k = torch.from_numpy(K).to(ins.device)
            _lib.convolve_beam(d_in, d_out, nx, ny, nc, d_k, K.shape[0], K.shape[1])
        out = d_out.cpu().numpy().astype(np.float64)
        dc._array = out[..., np.newaxis] if dc.stokes_axis else out
        dc.drop_pad()
        if dc._array_unit == "Jy/pix2":
            dc._array = dc._array / dc.px_size**2
        dc._array = dc._array * self.beam.area  # Jy/arcsec^2 -> Jy/beam
        dc._array_unit = "Jy/beam"
        if not self.quiet:
            a = dc._array
            print("Beam convolved.",
                  f"  Data cube RMS after beam convolution: {np.std(a):.2e} Jy / beam",
                  f"  Maximum pixel: {a.max():.2e} Jy / beam",
                  f"  Median non-zero pixel: {np.median(a[a > 0]) if (a > 0).any() else 0.0:.2e}"
                  " Jy / beam", sep="\n")

    def add_noise(self) -> None:
        """Add the noise model's realisation to the cube (martini.py:903-937):
        Jy/beam converted with the beam area to Jy/arcsec^2, and on to Jy/pix^2
        when the source is not inserted yet."""
        if self.noise is None:
            warnings.warn("Skipping noise, no noise object provided to Martini.")
            return
        if self.beam is None:
            warnings.warn("Skipping noise, no beam object (required to estimate "
                          "post-convolution noise level) provided to Martini.")
            return
        dc = self._datacube
        noise = np.asarray(self.noise.generate(dc, self.beam), float) / self.beam.area
        if dc._array_unit == "Jy/pix2":
            noise = noise * dc.px_size**2
        elif dc._array_unit != "Jy/arcsec2":
            raise ValueError(f"cannot add noise to a cube in {dc._array_unit}: add it before "
                             "the beam convolution")
        dc._array = dc._array + noise
        if not self.quiet:
            print("Noise added.",
                  f"  Noise cube RMS: {np.std(noise):.2e} (before beam convolution).",
                  "  Data cube RMS after noise addition (before beam convolution): "
                  f"{np.std(dc._array):.2e}", sep="\n")

    def write_fits(self, filename: str, overwrite: bool = True, obj_name: str = "MOCK",
                   channels=None, dtype=None) -> None:
        """Write the cube to a FITS file (martini.py:939-1055): '.fits' is
        appended unless present, the pad is dropped, the header holds the
        reference's cards, and the data unit (FITS axis order, big-endian,
        float64 unless ``dtype`` is float32) is made on the GPU."""
        from .insertion import default_inserter

        if channels is not None:
            warnings.warn(DeprecationWarning(
                "`channels` argument to `write_fits` ignored, channels and their units now "
                "fixed at DataCube initialization."))
        bitpix = fits.bitpix_for(dtype)
        dc = self._datacube
        dc.drop_pad()
        filename = fits.fits_filename(filename)
        ins = default_inserter()
        torch = ins.torch
        arr = dc._array[..., 0] if dc.stokes_axis else dc._array
        nx, ny, nc = arr.shape
        d_cube = torch.from_numpy(np.ascontiguousarray(arr, np.float64)).to(ins.device)
        d_out = torch.empty(nx * ny * nc * (-bitpix // 8), dtype=torch.uint8, device=ins.device)
        d_range = torch.empty(2, dtype=torch.float64, device=ins.device)
        _lib.cube_to_fits(d_cube, nx, ny, nc, 0, 0, bitpix, d_out, d_range)
        lo, hi = d_range.cpu().tolist()
        data = d_out.cpu().numpy().tobytes()
        cards = fits.cube_cards(dc, self.beam, obj_name, bitpix, lo, hi)
        fits.write_hdu(filename, cards, data, overwrite)

    def write_beam_fits(self, filename: str, overwrite: bool = True, channels=None) -> None:
        """Write the beam image to a FITS file (martini.py:1057-1152)."""
        if channels is not None:
            warnings.warn(DeprecationWarning(
                "`channels` argument to `write_fits` ignored, channels and their units now "
                "fixed at DataCube initialization."))
        if self.beam is None:
            raise ValueError("Martini.write_beam_fits: Called with beam set to 'None'.")
        filename = filename if filename[-5:] == ".fits" else filename + ".fits"
        k = np.asarray(self.beam.kernel, np.float64)
        # the reference writes kernel[..., newaxis].T: FITS axis 1 runs along kernel axis 0
        data = np.ascontiguousarray(k[..., np.newaxis].T).astype(">f8").tobytes()
        fits.write_hdu(filename, fits.beam_cards(self._datacube, self.beam, -64), data, True)

    def write_hdf5(self, filename: str, overwrite: bool = True, memmap: bool = False,
                   compact: bool = False, channels=None):
        """Write the cube and beam to an HDF5 file (martini.py:1154-1366): the
        pad is dropped and '.hdf5' appended unless present; the datasets and
        attributes are the reference's, in its order.  Without ``compact`` the
        world coordinates of every voxel centre (RA, Dec, channel_mids) and
        vertex (*_vertices) and the four channel arrays are written too.  With
        ``memmap`` the file is built in memory and returned open
        (``hdf5.File``); otherwise it is closed and None returned."""
        if channels is not None:
            warnings.warn(DeprecationWarning(
                "`channels` argument to `write_fits` ignored, channels and their units now "
                "fixed at DataCube initialization."))
        dc = self._datacube
        dc.drop_pad()
        filename = filename if filename[-5:] == ".hdf5" else filename + ".hdf5"
        ctype3, cunit3 = fits._spectral_axis(dc)
        crpix, cdelt, crval = dc.crpix, dc.cdelt, dc.crval
        f = hdf5.create(filename, overwrite, memmap)
        try:
            f.create_dataset("FluxCube", np.asarray(dc._array, np.float64).squeeze())
            if not compact:
                nx, ny, nc = dc.n_px_x, dc.n_px_y, dc.n_channels

                def grids(offset, n):
                    # the sky axes do not depend on the channel, nor the spectral on the pixel:
                    # broadcast, then materialised one dataset at a time
                    x = (np.arange(n[0]) + offset)[:, None, None]
                    y = (np.arange(n[1]) + offset)[None, :, None]
                    z = (np.arange(n[2]) + offset)[None, None, :]
                    return (np.broadcast_to(g, n) for g in dc.pixel_to_world(x, y, z))

                for names, n, off in ((("RA", "Dec", "channel_mids"), (nx, ny, nc), 0.0),
                                      (("RA_vertices", "Dec_vertices", "channel_vertices"),
                                       (nx + 1, ny + 1, nc + 1), -0.5)):
                    for name, grid, unit in zip(names, grids(off, n), ("deg", "deg", cunit3)):
                        f.create_dataset(name, grid, {"Unit": unit})
                for name, unit in (("velocity_channel_mids", "m / s"),
                                   ("velocity_channel_edges", "m / s"),
                                   ("frequency_channel_mids", "Hz"),
                                   ("frequency_channel_edges", "Hz")):
                    f.create_dataset(name, getattr(dc, name), {"Unit": unit})
            axes = {"RA": (1, "RA---TAN", "deg"), "Dec": (2, "DEC--TAN", "deg"),
                    "V": (3, ctype3, cunit3)}
            cube_attrs = {"AxisOrder": "(RA,Dec,Channels)", "FluxCubeUnit": _UNIT_STR[dc._array_unit]}
            for ax, (i, ctype, cunit) in axes.items():
                cube_attrs.update({f"delta{ax}_in_{ax}Unit": float(cdelt[i - 1]),
                                   f"{ax}0_in_px": float(crpix[i - 1]) - 1.0,
                                   f"{ax}0_in_{ax}Unit": float(crval[i - 1]),
                                   f"{ax}Unit": cunit, f"{ax}ProjType": ctype})
            cube_attrs["SpecSys"] = dc.specsys.upper()
            beam = self.beam
            beam_deg = {}
            if beam is not None:
                beam_deg = {"BeamPA": float(beam.bpa), "BeamMajor_in_deg": float(beam.bmaj) / 3600.0,
                            "BeamMinor_in_deg": float(beam.bmin) / 3600.0}
            stamp = {"DateCreated": hdf5.now_string(), "MartiniVersion": f"martini_amd {__version__}",
                     "AstropyVersion": "not used"}
            f.set_attrs("FluxCube", {**cube_attrs, **beam_deg, **stamp})
            if beam is not None:
                if beam.kernel is None:
                    raise ValueError("Martini.write_hdf5: Called with beam present but beam kernel"
                                     " uninitialized.")
                k = np.asarray(beam.kernel, np.float64)
                battrs = {"BeamUnit": ""}  # the beam image is dimensionless (beams.py:145-157)
                for ax, (i, ctype, cunit) in axes.items():
                    battrs.update({f"delta{ax}_in_{ax}Unit": float(cdelt[i - 1]),
                                   f"{ax}0_in_px": k.shape[i - 1] // 2 if i < 3 else 0,
                                   f"{ax}0_in_{ax}Unit": float(crval[i - 1]),
                                   f"{ax}Unit": cunit, f"{ax}ProjType": ctype})
                f.create_dataset("Beam", k[..., np.newaxis], {**battrs, **beam_deg, **stamp})
        except BaseException:
            f.close()
            raise
        if memmap:
            return f
        f.close()
        return None

    def insert_source_in_cube(self, skip_validation: bool = False, progressbar=None,
                              ncpu: int = 1, method: str = "plain") -> None:
        """Populate the cube with the source's flux (martini.py:827-861).

        ``method``: ``'plain'`` (FP32 tile gather) or ``'mfma'`` (BF16x3 MFMA).
        """
        self._insert_source_in_cube(skip_validation=skip_validation, progressbar=progressbar,
                                    ncpu=ncpu, method=method)


class GlobalProfile(_BaseMartini):
    """The spatially integrated spectrum of a source (martini.py:1369-1596).

    Every particle in the spectral band contributes its whole line: a 1 x 1
    pixel cube of 1 degree pixels (datacube.py:889-940), a Dirac kernel of
    unbounded range, no spatial pruning, and every particle placed on the one
    pixel.  The insertion itself is the same GPU path as ``Martini``'s.
    """

    def __init__(self, *, source, spectral_model, n_channels: int, channel_width: float,
                 spectral_centre: float = 0.0, quiet: bool = False, channels=None,
                 freq_channels: bool = False) -> None:
        from .datacube import DataCube
        from .sph_kernels import DiracDeltaKernel

        if channels is not None:
            warnings.warn(DeprecationWarning(
                "The `channels` argument to `GlobalProfile.__init__` is deprecated and has "
                "been ignored."))
        if source.hsm_g is None:
            source.hsm_g = np.zeros(source.npart)
        dc = DataCube(n_px_x=1, n_px_y=1, n_channels=n_channels, px_size=3600.0,
                      channel_width=channel_width, spectral_centre=spectral_centre,
                      freq_channels=freq_channels)
        super().__init__(source=source, datacube=dc, beam=None, noise=None,
                         sph_kernel=DiracDeltaKernel(size_in_fwhm=np.inf),
                         spectral_model=spectral_model, quiet=quiet,
                         _prune_kwargs=dict(spatial=False, spectral=True, mass=True,
                                            obj_type_str="spectrum"))
        self.source.pixcoords[:2] = 0

    def insert_source_in_spectrum(self, method: str = "plain") -> None:
        """Fill the spectrum (Jy per channel); martini.py:1550-1596."""
        super()._insert_source_in_cube(skip_validation=True, progressbar=False, ncpu=1,
                                       quiet=True, method=method)
        # Jy/arcsec^2 over the one pixel -> Jy
        self._spectrum = np.asarray(self._datacube._array).squeeze() * self._datacube.px_size**2
        if not self.quiet:
            flux = self._spectrum.sum()
            mass = (2.36e5 * self.source.distance**2 * flux
                    * np.abs(np.diff(self._datacube.velocity_channel_edges)).mean() / 1e3)
            print("Source inserted.", f"  Flux density in spectrum: {flux:.2e} Jy",
                  f"  Mass in spectrum (assuming distance {self.source.distance:.2f} Mpc): "
                  f"{mass:.2e} solMass",
                  f"    [{mass / self.source.input_mass * 100:.0f}% of initial source mass]",
                  sep="\n")

    @property
    def spectrum(self) -> np.ndarray:
        """The spectrum in Jy, evaluated on first access."""
        if not hasattr(self, "_spectrum"):
            self.insert_source_in_spectrum()
        return self._spectrum

    @property
    def channel_edges(self):
        return self._datacube.channel_edges

    @property
    def channel_mids(self):
        return self._datacube.channel_mids

    @property
    def velocity_channel_edges(self):
        return self._datacube.velocity_channel_edges

    @property
    def velocity_channel_mids(self):
        return self._datacube.velocity_channel_mids

    @property
    def frequency_channel_edges(self):
        return self._datacube.frequency_channel_edges

    @property
    def frequency_channel_mids(self):
        return self._datacube.frequency_channel_mids
