"""ctypes front end of ``cube_oracle.c`` (test infrastructure only).

``cube(P)`` returns the FP64 cube (Jy/pix^2, padded shape) for a
``martini_oracle.Prepared`` set of particles, with the same arithmetic as
``martini_oracle.cube_pixel_loop`` but particle-ordered and threaded.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import martini_oracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcube_oracle.so")
_lib = None

KIND_CODES = {"cubic": 0, "wc2": 1, "wc6": 2, "quartic": 3, "gauss": 4, "dirac": 5}


def build(force: bool = False) -> str:
    """Compile the C oracle with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        L.oracle_cube.restype = ctypes.c_int
        L.oracle_cube.argtypes = [ctypes.c_int64, dp, dp, dp, dp, ip, dp, dp, dp, dp, dp, dp,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, dp]
        L.oracle_weights.restype = None
        L.oracle_weights.argtypes = [ctypes.c_int64, ctypes.c_int, dp, dp, dp, ctypes.c_double,
                                     ctypes.c_double, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _per_particle_kernels(P: O.Prepared):
    """Concrete kernel code, Gaussian truncation and norm per particle."""
    n = P.h_px.size
    kinds = np.zeros(n, np.int32)
    gt = np.zeros(n)
    gn = np.ones(n)
    if isinstance(P.kernel, O.AdaptiveSpec):
        ks = P.kernel.kernels
        use = np.where(P.kernel_indices < 0, 0, P.kernel_indices)
    else:
        ks = (P.kernel,)
        use = np.zeros(n, int)
    for ik, K in enumerate(ks):
        m = use == ik
        kinds[m] = KIND_CODES[K.name]
        if K.name == "gauss":
            gt[m] = K.truncate
            gn[m] = K.norm
    return kinds, gt, gn


def cube(P: O.Prepared, nthreads: int | None = None) -> np.ndarray:
    """FP64 cube in Jy/pix^2 for prepared particles (padded shape)."""
    nx, ny, nc = P.cube.shape
    kinds, gt, gn = _per_particle_kernels(P)
    x = np.ascontiguousarray(P.pixcoords[0], float)
    y = np.ascontiguousarray(P.pixcoords[1], float)
    h_eff = np.ascontiguousarray(P.h_px * P.rescale, float)
    r = np.ascontiguousarray(P.sm_ranges, float)
    v = np.ascontiguousarray(P.vlos_kms, float)
    sig = np.ascontiguousarray(P.sigma_kms, float)
    amp = np.ascontiguousarray(P.mHI * P.distance_mpc ** -2.0, float)
    edges = np.ascontiguousarray(P.cube.velocity_channel_edges() / 1000.0, float)
    out = np.zeros((nx, ny, nc))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    rc = lib().oracle_cube(x.size, _dp(x), _dp(y), _dp(h_eff), _dp(r),
                           kinds.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(gt), _dp(gn),
                           _dp(v), _dp(sig), _dp(amp), _dp(edges), nx, ny, nc,
                           1 if P.spectrum == "dirac" else 0, int(nthreads), _dp(out))
    if rc != 0:
        raise MemoryError("oracle_cube could not allocate its per-thread cubes")
    return out


def weights(kind: str, dx, dy, h, truncate: float = 0.0) -> np.ndarray:
    """C oracle kernel integral at (dx, dy, h) arrays."""
    dx = np.ascontiguousarray(dx, float)
    dy = np.ascontiguousarray(np.broadcast_to(dy, dx.shape), float)
    h = np.ascontiguousarray(np.broadcast_to(h, dx.shape), float)
    out = np.empty_like(dx)
    norm = O.gauss_norm(truncate) if kind == "gauss" else 1.0
    lib().oracle_weights(dx.size, KIND_CODES[kind], _dp(dx), _dp(dy), _dp(h), float(truncate),
                         float(norm), _dp(out))
    return out
