"""Test helpers: run a workload through martini_amd and through the oracle."""

from __future__ import annotations

import numpy as np

from martini_amd import Martini, synthetic
from oracle import cube_c
from oracle import martini_oracle as O

VOXEL_TOL = 1e-5  # |gpu - oracle| / max(oracle), per voxel
FLUX_TOL = 1e-6  # |sum gpu - sum oracle| / sum oracle


def oracle_kernel(name, truncate=None):
    if name == "gauss":
        return O.adaptive(O.KernelSpec("gauss", truncate=truncate or 3.0))
    return O.adaptive(O.KernelSpec(name))


def oracle_cube(w: synthetic.Workload, n=None, kernel=None, spectrum="gauss", sigma="thermal"):
    """FP64 cube in Jy/arcsec^2 for a workload (optionally fewer particles)."""
    a = w.arrays(n)
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys)
    K = kernel if kernel is not None else oracle_kernel(w.kernel)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                  kernel=K, rotation=w.rotation.as_matrix(), spectrum=spectrum, sigma=sigma)
    return cube_c.cube(P) / w.px_size**2, P


def gpu_cube(w: synthetic.Workload, n=None, method="plain", kernel=None, spectrum=None):
    src, dc, k, s = synthetic.build(w, n, kernel=kernel, spectrum=spectrum)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    m.insert_source_in_cube(method=method)
    return m.datacube._array, m


def parity(gpu, ref):
    peak = np.abs(ref).max()
    voxel = np.abs(gpu - ref).max() / peak
    flux = abs(gpu.sum() - ref.sum()) / abs(ref.sum())
    return voxel, flux
