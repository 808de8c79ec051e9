"""FP64 source setup on the GPU (``mi_prepare``).

``observation`` reduces the plugin objects -- a ``DataCube``, an SPH kernel,
a spectral model -- and a source placement (distance, peculiar velocity,
Hubble h, ra/dec, rotation; the ``SPHSource`` keywords of
martini/sources/sph_source.py:171-263) to the ``mi_observation`` struct.
``DeviceSource`` holds a source's raw particle arrays in HBM, and
``prepare`` runs the setup that ``Martini.__init__`` does on the host
(martini/martini.py:118-151) on the device: sky and pixel coordinates,
smoothing lengths, adaptive kernel choice, pruning, record packing.  The
records then go to ``mi_insert`` / ``mi_bin`` + ``mi_gather`` unchanged.
"""

from __future__ import annotations

import numpy as np
from scipy.spatial.transform import Rotation

from . import _lib
from .sph_kernels import DiracDeltaKernel, _AdaptiveKernel, _GaussianKernel, _MinSizeKernel


def _tiers(sph_kernel):
    kernels = sph_kernel.kernels if isinstance(sph_kernel, _AdaptiveKernel) else (sph_kernel,)
    if len(kernels) > 3:
        raise ValueError("at most 3 kernel tiers")
    rows, gauss = [], []
    for K in kernels:
        kind = getattr(K, "_mi_kind", -1)
        if kind < 0:
            raise NotImplementedError(f"{type(K).__name__} has no device implementation")
        gidx = 0
        if isinstance(K, _GaussianKernel):
            key = (float(K.truncate), float(K.norm))
            if key not in gauss:
                gauss.append(key)
            gidx = gauss.index(key)
        if isinstance(K, DiracDeltaKernel):
            lo, hi = -np.inf, float(K.max_valid_size)
        elif isinstance(K, _MinSizeKernel):
            lo, hi = float(K.min_valid_size), np.inf
        else:
            lo, hi = -np.inf, np.inf
        rows.append((kind, gidx, float(K._rescale), float(K.size_in_fwhm), lo, hi))
    return rows, gauss


def observation(*, distance: float, datacube, sph_kernel, spectral_model, vpeculiar: float = 0.0,
                h: float = 0.7, ra: float = 0.0, dec: float = 0.0, rotation=None,
                out_scale: float | None = None, prune: bool = True):
    """(mi_observation, mi_cube_desc) for one source seen into ``datacube``.

    ``rotation``: scipy Rotation or 3x3 matrix applied to the raw positions and
    velocities (SPHSource's ``rotation``).  ``out_scale`` defaults to
    1/px_size^2, which gives Jy/arcsec^2 as ``insert_source_in_cube`` does.
    """
    o = _lib.Observation()
    rot = np.eye(3) if rotation is None else (
        rotation.as_matrix() if hasattr(rotation, "as_matrix") else np.asarray(rotation, float))
    rar, decr = np.deg2rad(ra), np.deg2rad(dec)
    sky = (Rotation.from_euler("z", rar) * Rotation.from_euler("y", -decr)).as_matrix()
    point = np.array([np.cos(decr) * np.cos(rar), np.cos(decr) * np.sin(rar), np.sin(decr)])
    o.rotation[:] = list(np.asarray(rot, float).ravel())
    o.sky_rotation[:] = list(sky.ravel())
    o.pointing[:] = list(point)
    o.distance_mpc, o.vpeculiar_kms, o.hubble_h = float(distance), float(vpeculiar), float(h)
    dc = datacube
    o.cube_ra_deg, o.cube_dec_deg, o.px_size_arcsec = dc.ra, dc.dec, dc.px_size
    o.crpix[:] = list(dc.crpix)
    o.cdelt[:] = list(dc.cdelt)
    o.crval2 = float(dc.crval[2])
    e = np.asarray(dc.velocity_channel_edges, float)
    o.e0_ms = float(e[0])
    o.de_ms = float((e[-1] - e[0]) / dc.n_channels)
    o.dv_max_kms = float(np.max(np.abs(np.diff(e))) / 1e3)
    kind = spectral_model._mi_kind
    if kind == _lib.SPECTRUM_TABLE:
        raise NotImplementedError(f"{type(spectral_model).__name__}: device setup covers the "
                                  "Gaussian and Dirac lines; Martini(...) evaluates other "
                                  "spectral models on the host")
    o.spectrum = kind
    if kind == _lib.SPECTRUM_GAUSS:
        mode = spectral_model.sigma_mode
        o.sigma_kms = -1.0 if isinstance(mode, str) else float(mode)
        if isinstance(mode, str) and mode != "thermal":
            raise ValueError("sigma must be a velocity in km/s or 'thermal'")
    else:
        o.sigma_kms = 0.0
    o.out_scale = float(1.0 / dc.px_size**2 if out_scale is None else out_scale)
    rows, gauss = _tiers(sph_kernel)
    if len(gauss) > 4:
        raise ValueError("at most 4 distinct Gaussian truncations per insertion")
    o.n_tiers = len(rows) if isinstance(sph_kernel, _AdaptiveKernel) else 1
    for k in range(3):
        kd, gi, resc, size, lo, hi = rows[min(k, len(rows) - 1)]
        o.tier_kind[k], o.tier_gidx[k] = kd, gi
        o.tier_rescale[k], o.tier_size[k], o.tier_min[k], o.tier_max[k] = resc, size, lo, hi
    o.nx, o.ny, o.padx, o.pady, o.nc = dc.n_px_x, dc.n_px_y, dc.padx, dc.pady, dc.n_channels
    o.freq_channels = int(dc._freq_channel_mode)
    o.prune = int(bool(prune))
    desc = _lib.make_desc(dc.n_px_x + 2 * dc.padx, dc.n_px_y + 2 * dc.pady, dc.n_channels, kind,
                          gauss)
    return o, desc


class DeviceSource:
    """Raw FP64 particle arrays of one source in HBM (positions kpc, velocities
    km/s, masses Msun, smoothing lengths kpc, temperatures K)."""

    def __init__(self, xyz, vxyz, mHI, hsm, T=None, device=None) -> None:
        import torch

        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device

        def f64(a, shape=None):
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, float))
            t = t.to(device=dev, dtype=torch.float64).contiguous()
            return t if shape is None else t.reshape(shape)

        self.xyz = f64(xyz, (-1, 3))
        self.n = int(self.xyz.shape[0])
        self.vxyz = f64(vxyz, (self.n, 3))
        self.mHI = f64(mHI, (self.n,))
        self.hsm = f64(hsm, (self.n,))
        self.T = None if T is None else f64(T, (self.n,))
        self.device = dev

    @property
    def nbytes(self) -> int:
        return self.n * 8 * (9 + (self.T is not None))


def prepare(src: DeviceSource, obs, datacube=None, stream=None, d_edges=None, out=None, ws=None):
    """Runs mi_prepare; returns (d_records, d_counts) with d_counts a 4-element
    int64 device tensor (kept, cell pairs, flags, mask pixels).  Asynchronous.
    The kept records are in input order; for the FP32 form, inputs that come
    spatially ordered are best inserted from
    ``insertion.shuffle_records(d_records[: kept * 48])``, the order host
    records take at upload."""
    import torch

    dev = src.device
    if d_edges is None and obs.spectrum == _lib.SPECTRUM_DIRAC:
        if datacube is None:
            raise ValueError("the Dirac spectrum needs the datacube's channel edges")
        d_edges = torch.from_numpy(np.asarray(datacube.velocity_channel_edges, float)).to(dev)
    if out is None:
        out = torch.empty(max(src.n, 1) * 48, dtype=torch.uint8, device=dev)
    counts = torch.empty(4, dtype=torch.int64, device=dev)
    if ws is None:
        ws = torch.empty(max(_lib.prepare_workspace_bytes(src.n), 256), dtype=torch.uint8,
                         device=dev)
    _lib.prepare(src.xyz, src.vxyz, src.mHI, src.hsm, src.T, d_edges, src.n, obs, out, counts, ws,
                 stream)
    return out, counts


def check_flags(flags: int, skip_validation: bool = False) -> None:
    """Raise as the host path would for the flags mi_prepare reported."""
    if flags & _lib.PREP_COORD_LIMIT:
        raise ValueError("pixel coordinates beyond 2^21 cannot be represented")
    if (flags & _lib.PREP_INVALID_KERNEL) and not skip_validation:
        raise RuntimeError("Some particles fail SPH kernel validation; pass skip_validation=True "
                           "to insert them anyway.")


__all__ = ["observation", "DeviceSource", "prepare", "check_flags"]
