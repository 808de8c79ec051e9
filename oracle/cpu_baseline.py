"""The reference's CPU insertion algorithm, restated as it runs, for timing.

TEST/BASELINE INFRASTRUCTURE ONLY (bench.py's ``cpu_baseline`` leg).

Follows martini/martini.py:285-362 and :243-283: precompute the N x Nc FP64
spectra (init_spectra, spectral_models.py:63-147), then for every pixel of
the padded cube scan all particles with the square mask
|i - x| <= r and |j - y| <= r, evaluate the kernel weights of the masked
particles, gather their spectra rows, multiply and sum; pixels are mapped
serially or through ``ThreadPool(ncpu).imap`` (order preserving), as the
reference does (martini.py:344-362).  astropy's unit bookkeeping is left out,
so this is a faster-than-reference baseline.
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

# Host threads a job may use on a shared GPU host: the CPU share of one GPU.
HOST_THREAD_SHARE = 16


def host_threads() -> int:
    """Threads for an all-core run: the CPUs this process may run on, at most
    the per-GPU host share."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(n, HOST_THREAD_SHARE))


def evaluate_pixel(P: O.Prepared, S: np.ndarray, ij, rows=None) -> np.ndarray:
    """_evaluate_pixel_spectrum (martini/martini.py:243-283).  ``rows`` maps a
    particle index to its row of ``S`` when S holds a subset of the particles."""
    i, j = ij
    x, y = P.pixcoords[0], P.pixcoords[1]
    m = (np.abs(i - x) <= P.sm_ranges) & (np.abs(j - y) <= P.sm_ranges)
    w = O.kernel_weights(P, x[m] - i, y[m] - j, m)
    tmp = S[m] if rows is None else S[rows[m]]
    np.multiply(tmp, w[:, None], out=tmp)
    return tmp.sum(axis=-2)


def insert(P: O.Prepared, S: np.ndarray, pixels, ncpu: int = 1, rows=None) -> np.ndarray:
    """The pixel loop over ``pixels`` (serial or ThreadPool imap)."""
    if ncpu == 1:
        out = [evaluate_pixel(P, S, ij, rows) for ij in pixels]
    else:
        with ThreadPool(processes=ncpu) as pool:
            out = list(pool.imap(lambda ij: evaluate_pixel(P, S, ij, rows), pixels))
    return np.array(out)


def updates(P: O.Prepared) -> int:
    """Particle-voxel updates: Nc * sum_p n_x(p) n_y(p) over the padded cube,
    with the reference's mask |i - x| <= r."""
    nx, ny, nc = P.cube.shape
    x, y, r = P.pixcoords[0], P.pixcoords[1], P.sm_ranges
    nxp = np.clip(np.minimum(nx - 1, np.floor(x + r)) - np.maximum(0, np.ceil(x - r)) + 1, 0, None)
    nyp = np.clip(np.minimum(ny - 1, np.floor(y + r)) - np.maximum(0, np.ceil(y - r)) + 1, 0, None)
    return int(nc * np.sum(nxp * nyp))


def _subset(P: O.Prepared, sel) -> O.Prepared:
    Q = O.Prepared()
    for k, v in vars(P).items():
        setattr(Q, k, v[..., sel] if isinstance(v, np.ndarray) and v.shape[-1:] == P.h_px.shape
                else v)
    return Q


def run_full(P: O.Prepared, ncpu: int = 1) -> dict:
    """The whole insertion: init_spectra, then every pixel of the padded cube."""
    nx, ny, _ = P.cube.shape
    t0 = time.perf_counter()
    S = O.particle_spectra(P)
    t_spec = time.perf_counter() - t0
    pixels = [(i, j) for i in range(nx) for j in range(ny)]
    t0 = time.perf_counter()
    insert(P, S, pixels, ncpu)
    t_loop = time.perf_counter() - t0
    U = updates(P)
    return {
        "value": U / (t_spec + t_loop),
        "unit": "particle-voxel updates/s",
        "cores": int(ncpu),
        "kind": "port",
        "sample": (f"full run: init_spectra of {P.h_px.size} particles ({t_spec:.2f} s) and all "
                   f"{len(pixels)} pixels ({t_loop:.2f} s); "
                   + ("serial loop" if ncpu == 1 else f"ThreadPool({ncpu})")),
        "seconds": t_spec + t_loop,
    }


def measure(P: O.Prepared, budget_s: float = 15.0, ncpu: int | None = None,
            spectra_budget_s: float = 5.0) -> dict:
    """Time the insertion on a deterministic sample and extrapolate linearly.

    init_spectra is O(N Nc): it is timed on the first particles (all of them
    when that fits ``spectra_budget_s``) and scaled by N.  The pixel loop's
    per-pixel cost is dominated by the O(N) mask, the same for every pixel: it
    is timed on pixels of every k-th row of the padded cube (every m-th pixel of
    those rows), sized to ``budget_s``, with the spectra of the particles those
    rows mask, and scaled by the number of pixels."""
    if ncpu is None:
        ncpu = host_threads()
    nx, ny, nc = P.cube.shape
    n = P.h_px.size
    # init_spectra: a probe, then as many particles as the budget allows
    probe = _subset(P, slice(0, min(n, 2000)))
    t0 = time.perf_counter()
    O.particle_spectra(probe)
    per_part = (time.perf_counter() - t0) / probe.h_px.size
    n_spec = int(min(n, max(probe.h_px.size, spectra_budget_s / max(per_part, 1e-12))))
    sub = P if n_spec == n else _subset(P, slice(0, n_spec))
    t0 = time.perf_counter()
    O.particle_spectra(sub)
    t_spec = (time.perf_counter() - t0) * n / n_spec
    # pixel loop: a probe of a few pixels gives the per-pixel cost
    x, r = P.pixcoords[0], P.sm_ranges
    probe_px = [(nx // 2, j) for j in range(0, ny, max(1, ny // (2 * ncpu)))][: 2 * ncpu]
    need = np.flatnonzero(np.abs(nx // 2 - x) <= r)
    rows = np.full(n, -1, np.int64)
    rows[need] = np.arange(need.size)
    S = O.particle_spectra(_subset(P, need))
    t0 = time.perf_counter()
    insert(P, S, probe_px, ncpu, rows)
    per_px = (time.perf_counter() - t0) / len(probe_px)
    n_px = int(min(nx * ny, max(len(probe_px), budget_s / max(per_px, 1e-9))))
    n_rows = min(nx, -(-n_px // ny))
    stride = max(1, nx // n_rows)
    sample_rows = list(range(stride // 2, nx, stride))[:n_rows]
    col_stride = max(1, (n_rows * ny) // n_px)
    sel = np.zeros(n, bool)
    for i in sample_rows:
        sel |= np.abs(i - x) <= r
    need = np.flatnonzero(sel)
    rows[:] = -1
    rows[need] = np.arange(need.size)
    S = O.particle_spectra(_subset(P, need))
    pixels = [(i, j) for i in sample_rows for j in range(ny)][::col_stride]
    t0 = time.perf_counter()
    insert(P, S, pixels, ncpu, rows)
    t_loop = time.perf_counter() - t0
    t_total = t_spec + t_loop * (nx * ny) / len(pixels)
    U = updates(P)
    return {
        "value": U / t_total,
        "unit": "particle-voxel updates/s",
        "cores": int(ncpu),
        "kind": "port",
        "sample": (f"init_spectra timed on {n_spec} of {n} particles ({t_spec:.2f} s scaled to all) "
                   f"+ pixel loop on {len(pixels)} px of {len(sample_rows)} of {nx} pixel rows "
                   f"(one pixel in {col_stride} of each, {t_loop:.2f} s), extrapolated linearly "
                   f"to {nx * ny} px; ThreadPool({ncpu})"),
        "seconds_extrapolated": t_total,
    }
