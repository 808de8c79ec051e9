"""The reference's CPU insertion algorithm, restated as it runs, for timing.

TEST/BASELINE INFRASTRUCTURE ONLY (bench.py's ``cpu_baseline`` leg).

Follows martini/martini.py:285-362 and :243-283: precompute the N x Nc FP64
spectra (init_spectra, spectral_models.py:63-147), then for every pixel of
the padded cube scan all particles with the square mask
|i - x| <= r and |j - y| <= r, evaluate the kernel weights of the masked
particles, gather their spectra rows, multiply and sum; pixels are mapped
serially or through ``ThreadPool(ncpu).imap`` (order preserving), as the
reference does.  astropy's unit bookkeeping is left out, so this is a
faster-than-reference baseline.
"""

from __future__ import annotations

import os
import time

import numpy as np

from . import martini_oracle as O

try:  # the reference uses multiprocess's ThreadPool; the stdlib one is the same pool
    from multiprocess.pool import ThreadPool
except ImportError:  # pragma: no cover
    from multiprocessing.pool import ThreadPool


def evaluate_pixel(P: O.Prepared, S: np.ndarray, ij) -> np.ndarray:
    """_evaluate_pixel_spectrum (martini/martini.py:243-283)."""
    i, j = ij
    x, y = P.pixcoords[0], P.pixcoords[1]
    m = (np.abs(i - x) <= P.sm_ranges) & (np.abs(j - y) <= P.sm_ranges)
    w = O.kernel_weights(P, x[m] - i, y[m] - j, m)
    tmp = S[m]
    np.multiply(tmp, w[:, None], out=tmp)
    return tmp.sum(axis=-2)


def insert(P: O.Prepared, S: np.ndarray, pixels, ncpu: int = 1) -> np.ndarray:
    """The pixel loop over ``pixels`` (serial or ThreadPool imap)."""
    if ncpu == 1:
        out = [evaluate_pixel(P, S, ij) for ij in pixels]
    else:
        with ThreadPool(processes=ncpu) as pool:
            out = list(pool.imap(lambda ij: evaluate_pixel(P, S, ij), pixels))
    return np.array(out)


def updates(P: O.Prepared) -> int:
    """Particle-voxel updates: Nc * sum_p n_x(p) n_y(p) over the padded cube,
    with the reference's mask |i - x| <= r."""
    nx, ny, nc = P.cube.shape
    x, y, r = P.pixcoords[0], P.pixcoords[1], P.sm_ranges
    nxp = np.clip(np.minimum(nx - 1, np.floor(x + r)) - np.maximum(0, np.ceil(x - r)) + 1, 0, None)
    nyp = np.clip(np.minimum(ny - 1, np.floor(y + r)) - np.maximum(0, np.ceil(y - r)) + 1, 0, None)
    return int(nc * np.sum(nxp * nyp))


def measure(P: O.Prepared, budget_s: float = 15.0, ncpu: int | None = None) -> dict:
    """Time init_spectra in full and the pixel loop on a deterministic sample
    of pixels sized to ~budget_s; extrapolate the loop linearly (its per-pixel
    cost is dominated by the O(N) mask, the same for every pixel)."""
    if ncpu is None:
        ncpu = min(16, os.cpu_count() or 1)
    nx, ny, nc = P.cube.shape
    t0 = time.perf_counter()
    S = O.particle_spectra(P)
    t_spec = time.perf_counter() - t0
    all_px = [(i, j) for i in range(nx) for j in range(ny)]
    probe = all_px[:: max(1, len(all_px) // (4 * ncpu))][: 4 * ncpu]
    t0 = time.perf_counter()
    insert(P, S, probe, ncpu)
    per_px = (time.perf_counter() - t0) / len(probe)
    n_sample = int(min(len(all_px), max(len(probe), budget_s / max(per_px, 1e-9))))
    stride = max(1, len(all_px) // n_sample)
    sample = all_px[::stride]
    t0 = time.perf_counter()
    insert(P, S, sample, ncpu)
    t_loop = time.perf_counter() - t0
    t_total = t_spec + t_loop * len(all_px) / len(sample)
    U = updates(P)
    return {
        "value": U / t_total,
        "unit": "particle-voxel updates/s",
        "cores": int(ncpu),
        "kind": "port",
        "sample": (f"init_spectra of all {P.h_px.size} particles ({t_spec:.2f} s) + pixel loop "
                   f"on one pixel in {stride} of {len(all_px)} ({len(sample)} px, {t_loop:.2f} s), "
                   f"loop extrapolated linearly; ThreadPool({ncpu})"),
        "seconds_extrapolated": t_total,
    }
