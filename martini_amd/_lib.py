"""ctypes binding of libmartini_mi355x.so (the C ABI in include/martini_mi355x.h).

The library is required: importing the compute entry points without it, or
calling them without a GPU, raises.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmartini_mi355x.so")

MI_OK = 0
MI_ERR_WORKSPACE = -2

KERNEL_CUBIC, KERNEL_WC2, KERNEL_WC6, KERNEL_QUARTIC, KERNEL_GAUSS, KERNEL_DIRAC = range(6)
KERNEL_TABLE = 6  # a kernel without a device form: weights evaluated by the host
SPECTRUM_GAUSS, SPECTRUM_DIRAC, SPECTRUM_TABLE = 0, 1, 2
ABI_VERSION = 3
METHOD_PLAIN_F32, METHOD_MFMA_BF16X3 = 0, 1
METHODS = {"plain": METHOD_PLAIN_F32, "fp32": METHOD_PLAIN_F32,
           "mfma": METHOD_MFMA_BF16X3, "bf16x3": METHOD_MFMA_BF16X3}

# struct mi_particle (48 bytes)
PARTICLE_DTYPE = np.dtype([
    ("ix", "<i4"), ("iy", "<i4"), ("fx", "<f4"), ("fy", "<f4"), ("h", "<f4"),
    ("r_kind", "<i4"), ("c_lo", "<i4"), ("c_hi", "<i4"), ("k0", "<i4"), ("fu", "<f4"),
    ("s", "<f4"), ("amp", "<f4"),
])
assert PARTICLE_DTYPE.itemsize == 48


class CubeDesc(ctypes.Structure):
    """struct mi_cube_desc."""

    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nc", ctypes.c_int32),
                ("spectrum", ctypes.c_int32), ("gauss_t", ctypes.c_float * 4),
                ("gauss_inv_norm", ctypes.c_float * 4), ("spectra", ctypes.c_void_p),
                ("spectra_ld", ctypes.c_int64), ("weights", ctypes.c_void_p),
                ("weights_n", ctypes.c_int64)]


class Observation(ctypes.Structure):
    """struct mi_observation (device FP64 setup of one source)."""

    _d3 = ctypes.c_double * 3
    _fields_ = [("rotation", ctypes.c_double * 9), ("sky_rotation", ctypes.c_double * 9),
                ("pointing", _d3), ("distance_mpc", ctypes.c_double),
                ("vpeculiar_kms", ctypes.c_double), ("hubble_h", ctypes.c_double),
                ("cube_ra_deg", ctypes.c_double), ("cube_dec_deg", ctypes.c_double),
                ("px_size_arcsec", ctypes.c_double), ("crpix", _d3), ("cdelt", _d3),
                ("crval2", ctypes.c_double), ("e0_ms", ctypes.c_double), ("de_ms", ctypes.c_double),
                ("dv_max_kms", ctypes.c_double), ("sigma_kms", ctypes.c_double),
                ("out_scale", ctypes.c_double), ("tier_rescale", _d3), ("tier_size", _d3),
                ("tier_min", _d3), ("tier_max", _d3), ("tier_kind", ctypes.c_int32 * 3),
                ("tier_gidx", ctypes.c_int32 * 3), ("n_tiers", ctypes.c_int32),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("padx", ctypes.c_int32),
                ("pady", ctypes.c_int32), ("nc", ctypes.c_int32), ("spectrum", ctypes.c_int32),
                ("freq_channels", ctypes.c_int32), ("prune", ctypes.c_int32),
                ("channel_offset", ctypes.c_int32), ("nc_total", ctypes.c_int32)]


PREP_INVALID_KERNEL = 1
PREP_COORD_LIMIT = 2


class MartiniLibError(RuntimeError):
    """A C-ABI call returned an error code."""


_lib = None


def lib() -> ctypes.CDLL:
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m martini_amd.build_native` "
            "(the insertion path has no CPU fallback)")
    # torch carries its own HIP runtime; loading torch first makes the library's
    # libamdhip64 dependency resolve to that same runtime, so device memory and
    # streams from torch are valid in the library's calls.
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - torch is the device-memory provider
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
    L.mi_abi_version.restype = ctypes.c_int
    if L.mi_abi_version() != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI version {L.mi_abi_version()}, expected "
                          f"{ABI_VERSION}: rebuild with python -m martini_amd.build_native")
    L.mi_strerror.restype = ctypes.c_char_p
    L.mi_strerror.argtypes = [ctypes.c_int]
    L.mi_count_pairs_host.restype = i64
    L.mi_count_pairs_host.argtypes = [vp, i64, ctypes.POINTER(CubeDesc)]
    L.mi_insert_workspace_size.restype = ctypes.c_int
    L.mi_insert_workspace_size.argtypes = [ctypes.POINTER(CubeDesc), i64, i64,
                                           ctypes.POINTER(ctypes.c_size_t)]
    L.mi_insert.restype = ctypes.c_int
    L.mi_insert.argtypes = [vp, i64, i64, ctypes.POINTER(CubeDesc), vp, vp, ctypes.c_size_t,
                            ctypes.c_int, vp]
    L.mi_insert_noalloc.restype = ctypes.c_int
    L.mi_insert_noalloc.argtypes = [vp, i64, i64, ctypes.POINTER(CubeDesc), vp, ctypes.c_int, vp]
    L.mi_bin.restype = ctypes.c_int
    L.mi_bin.argtypes = [vp, i64, i64, ctypes.POINTER(CubeDesc), vp, ctypes.c_size_t,
                         ctypes.c_int, vp]
    L.mi_gather.restype = ctypes.c_int
    L.mi_gather.argtypes = [vp, i64, i64, ctypes.POINTER(CubeDesc), vp, vp, ctypes.c_size_t,
                            ctypes.c_int, vp]
    L.mi_status.restype = ctypes.c_int
    L.mi_status.argtypes = [ctypes.POINTER(CubeDesc), i64, i64, vp,
                            ctypes.POINTER(ctypes.c_uint32 * 4), vp]
    L.mi_kernel_weights.restype = ctypes.c_int
    L.mi_kernel_weights.argtypes = [ctypes.c_int, f32, f32, vp, vp, vp, vp, i64, vp]
    L.mi_spectra.restype = ctypes.c_int
    L.mi_spectra.argtypes = [vp, i64, ctypes.POINTER(CubeDesc), vp, vp]
    L.mi_convolve_beam.restype = ctypes.c_int
    L.mi_convolve_beam.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp,
                                   ctypes.c_int32, ctypes.c_int32, vp]
    L.mi_cube_to_fits.restype = ctypes.c_int
    L.mi_cube_to_fits.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                  vp, vp, vp]
    L.mi_prepare_workspace_size.restype = ctypes.c_int
    L.mi_prepare_workspace_size.argtypes = [i64, ctypes.POINTER(ctypes.c_size_t)]
    L.mi_prepare.restype = ctypes.c_int
    L.mi_prepare.argtypes = [vp, vp, vp, vp, vp, vp, i64, ctypes.POINTER(Observation), vp, vp,
                             vp, ctypes.c_size_t, vp]
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != MI_OK:
        msg = lib().mi_strerror(rc).decode()
        raise MartiniLibError(f"{what} failed: {msg} (code {rc})")


def _torch_cuda():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("martini_amd's insertion path runs on the GPU and no GPU is visible")
    return torch


def stream_ptr(torch_stream=None):
    torch = _torch_cuda()
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def make_desc(nx, ny, nc, spectrum, gauss_table=(), spectra=None, spectra_ld=0) -> CubeDesc:
    """Cube descriptor; ``spectra`` is the device table (a tensor or a
    pointer) for SPECTRUM_TABLE, with row stride ``spectra_ld`` floats."""
    d = CubeDesc()
    d.nx, d.ny, d.nc, d.spectrum = int(nx), int(ny), int(nc), int(spectrum)
    d.spectra = spectra.data_ptr() if hasattr(spectra, "data_ptr") else spectra
    d.spectra_ld = int(spectra_ld)
    for k in range(4):
        d.gauss_t[k] = 0.0
        d.gauss_inv_norm[k] = 1.0
    for k, (t, norm) in enumerate(gauss_table):
        d.gauss_t[k] = float(t)
        d.gauss_inv_norm[k] = float(1.0 / norm)
    return d


def count_pairs(records: np.ndarray, desc: CubeDesc) -> int:
    rec = np.ascontiguousarray(records)
    n = lib().mi_count_pairs_host(rec.ctypes.data_as(ctypes.c_void_p), rec.size, ctypes.byref(desc))
    if n < 0:
        check(int(n), "mi_count_pairs_host")
    return int(n)


def workspace_bytes(desc: CubeDesc, n: int, n_pairs: int) -> int:
    out = ctypes.c_size_t(0)
    check(lib().mi_insert_workspace_size(ctypes.byref(desc), n, n_pairs, ctypes.byref(out)),
          "mi_insert_workspace_size")
    return int(out.value)


def insert(d_records, n: int, n_pairs: int, desc: CubeDesc, d_cube, d_ws, method="plain",
           stream=None) -> None:
    """cube += insertion of n device records (torch tensors on the GPU)."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    rc = lib().mi_insert(ctypes.c_void_p(d_records.data_ptr()), n, n_pairs, ctypes.byref(desc),
                         ctypes.c_void_p(d_cube.data_ptr()), ctypes.c_void_p(d_ws.data_ptr()),
                         d_ws.numel() * d_ws.element_size(), m, stream_ptr(stream))
    check(rc, "mi_insert")


SCRATCH_BYTES = 256 << 20  # MI_SCRATCH_BYTES
MI_ERR_SCRATCH = -6


def insert_noalloc(d_records, n: int, n_pairs: int, desc: CubeDesc, d_cube, method="plain",
                   stream=None) -> int:
    """mi_insert_noalloc: cube += insertion in the library's static scratch (no
    workspace, nothing allocated); returns the status code unchanged."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    return int(lib().mi_insert_noalloc(ctypes.c_void_p(d_records.data_ptr()), n, n_pairs,
                                       ctypes.byref(desc), ctypes.c_void_p(d_cube.data_ptr()), m,
                                       stream_ptr(stream)))


def bin(d_records, n: int, n_pairs: int, desc: CubeDesc, d_ws, stream=None,
        method: str = "plain") -> None:
    """Asynchronous binning stage (mi_bin), ordering the cells' particles for
    the gather ``method`` that follows."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    check(lib().mi_bin(ctypes.c_void_p(d_records.data_ptr()), n, n_pairs, ctypes.byref(desc),
                       ctypes.c_void_p(d_ws.data_ptr()), d_ws.numel() * d_ws.element_size(), m,
                       stream_ptr(stream)), "mi_bin")


def gather(d_records, n: int, n_pairs: int, desc: CubeDesc, d_cube, d_ws, method="plain",
           stream=None) -> None:
    """Asynchronous gather + reduction stage (mi_gather)."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    check(lib().mi_gather(ctypes.c_void_p(d_records.data_ptr()), n, n_pairs, ctypes.byref(desc),
                          ctypes.c_void_p(d_cube.data_ptr()), ctypes.c_void_p(d_ws.data_ptr()),
                          d_ws.numel() * d_ws.element_size(), m, stream_ptr(stream)), "mi_gather")


def status(desc: CubeDesc, n: int, n_pairs: int, d_ws, stream=None):
    """(pairs, overflow, items, split items) of the last binning; synchronises."""
    out = (ctypes.c_uint32 * 4)()
    rc = lib().mi_status(ctypes.byref(desc), n, n_pairs, ctypes.c_void_p(d_ws.data_ptr()),
                         ctypes.byref(out), stream_ptr(stream))
    if rc not in (MI_OK, MI_ERR_WORKSPACE):
        check(rc, "mi_status")
    return tuple(int(v) for v in out)


def kernel_weights(kind: int, dx, dy, h, gauss_t=0.0, gauss_norm=1.0) -> np.ndarray:
    """W(dx, dy, h) for one concrete kernel kind, evaluated on the GPU (FP32)."""
    torch = _torch_cuda()
    dx = np.ascontiguousarray(dx, np.float32).ravel()
    n = dx.size
    dy = np.ascontiguousarray(np.broadcast_to(np.asarray(dy, np.float32).ravel(), (n,)))
    h = np.ascontiguousarray(np.broadcast_to(np.asarray(h, np.float32).ravel(), (n,)))
    dev = torch.device("cuda", torch.cuda.current_device())
    tx, ty, th = (torch.from_numpy(a).to(dev) for a in (dx, dy, h))
    out = torch.empty(n, dtype=torch.float32, device=dev)
    rc = lib().mi_kernel_weights(int(kind), float(gauss_t), float(1.0 / gauss_norm),
                                 ctypes.c_void_p(tx.data_ptr()), ctypes.c_void_p(ty.data_ptr()),
                                 ctypes.c_void_p(th.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                 n, stream_ptr())
    check(rc, "mi_kernel_weights")
    return out.cpu().numpy()


def spectra(records: np.ndarray, desc: CubeDesc) -> np.ndarray:
    """N x Nc spectra (Jy, FP32) of host records, evaluated on the GPU."""
    torch = _torch_cuda()
    dev = torch.device("cuda", torch.cuda.current_device())
    rec = torch.from_numpy(np.ascontiguousarray(records).view(np.uint8)).to(dev)
    n = records.size
    out = torch.empty((n, desc.nc), dtype=torch.float32, device=dev)
    check(lib().mi_spectra(ctypes.c_void_p(rec.data_ptr()), n, ctypes.byref(desc),
                           ctypes.c_void_p(out.data_ptr()), stream_ptr()), "mi_spectra")
    return out.cpu().numpy()


def prepare_workspace_bytes(n: int) -> int:
    out = ctypes.c_size_t(0)
    check(lib().mi_prepare_workspace_size(int(n), ctypes.byref(out)), "mi_prepare_workspace_size")
    return int(out.value)


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def prepare(d_xyz, d_vxyz, d_mHI, d_hsm, d_T, d_edges, n: int, obs: Observation, d_out,
            d_counts, d_ws, stream=None) -> None:
    """Asynchronous device FP64 setup (mi_prepare); d_counts receives
    (kept, cell pairs, flags, mask pixels) as four uint64."""
    check(lib().mi_prepare(_ptr(d_xyz), _ptr(d_vxyz), _ptr(d_mHI), _ptr(d_hsm), _ptr(d_T),
                           _ptr(d_edges), int(n), ctypes.byref(obs), _ptr(d_out), _ptr(d_counts),
                           _ptr(d_ws), d_ws.numel() * d_ws.element_size(), stream_ptr(stream)),
          "mi_prepare")


def convolve_beam(d_in, d_out, nx: int, ny: int, nc: int, d_kernel, kx: int, ky: int,
                  stream=None) -> None:
    """Asynchronous beam convolution of a device cube (mi_convolve_beam)."""
    check(lib().mi_convolve_beam(_ptr(d_in), _ptr(d_out), int(nx), int(ny), int(nc),
                                 _ptr(d_kernel), int(kx), int(ky), stream_ptr(stream)),
          "mi_convolve_beam")


def cube_to_fits(d_cube, nx: int, ny: int, nc: int, padx: int, pady: int, bitpix: int, d_out,
                 d_range, stream=None) -> None:
    """Asynchronous FITS data unit of a device cube (mi_cube_to_fits): the
    unpadded region in FITS axis order, big-endian, and its min and max.
    d_cube is an FP32 or FP64 tensor; d_range two FP64 values."""
    bits = 8 * d_cube.element_size()
    check(lib().mi_cube_to_fits(_ptr(d_cube), bits, int(nx), int(ny), int(nc), int(padx),
                                int(pady), int(bitpix), _ptr(d_out), _ptr(d_range),
                                stream_ptr(stream)),
          "mi_cube_to_fits")
