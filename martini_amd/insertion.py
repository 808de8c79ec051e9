"""Device-side driver of the insertion path.

``Inserter`` owns a reusable workspace on one GPU and runs
bin -> gather -> reduce through the C ABI on the current torch stream.
Inputs are host records (``_pack.pack_particles``) or records already in HBM.
"""

from __future__ import annotations

import numpy as np

from . import _lib


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("martini_amd's insertion path runs on the GPU and no GPU is visible")
    return torch


class Inserter:
    """Keeps a workspace and runs insertions on one device."""

    def __init__(self, device=None) -> None:
        torch = _torch()
        self.torch = torch
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self._ws = None

    def workspace(self, nbytes: int):
        torch = self.torch
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = None
            self._ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=self.device)
        return self._ws

    def upload(self, records: np.ndarray, method="plain", nc=None):
        """Copy host records into HBM (uint8 tensor of n * 48 bytes), laid out
        for `method` and a cube of `nc` channels (``layout_records``)."""
        rec = np.ascontiguousarray(records)
        return layout_records(self.torch.from_numpy(rec.view(np.uint8)).to(self.device), method,
                              nc)

    def insert_device(self, d_records, n: int, n_pairs: int, desc, d_cube, method="plain",
                      stream=None) -> None:
        """cube += insertion of n device records; d_cube is a float32 tensor of
        nx*ny*nc elements."""
        nbytes = _lib.workspace_bytes(desc, n, n_pairs)
        ws = self.workspace(nbytes)
        _lib.insert(d_records, n, n_pairs, desc, d_cube, ws, method=method, stream=stream)

    def insert(self, records: np.ndarray, desc, method="plain", d_cube=None):
        """Insert host records; returns the device cube (float32, [nx][ny][nc])."""
        torch = self.torch
        n = int(records.size)
        d_spectra = None
        if desc.spectrum == _lib.SPECTRUM_TABLE:
            host = getattr(desc, "host_spectra", None)
            if host is None and not desc.spectra:
                raise ValueError("a table spectrum needs its spectra (host_spectra or a device "
                                 "pointer in the descriptor)")
            if host is not None:
                d_spectra = torch.from_numpy(np.ascontiguousarray(host, np.float32)).to(self.device)
                desc.spectra = d_spectra.data_ptr()
                desc.spectra_ld = host.shape[1]
        d_weights = None
        host_w = getattr(desc, "host_weights", None)
        if host_w is not None:
            d_weights = torch.from_numpy(np.ascontiguousarray(host_w, np.float32)).to(self.device)
            desc.weights = d_weights.data_ptr() if host_w.size else None
            desc.weights_n = int(host_w.size)
        n_pairs = _lib.count_pairs(records, desc)
        d_rec = self.upload(records, method, desc.nc)
        if d_cube is None:
            d_cube = torch.zeros(desc.nx * desc.ny * desc.nc, dtype=torch.float32, device=self.device)
        self.insert_device(d_rec, n, n_pairs, desc, d_cube, method=method)
        if d_spectra is not None or d_weights is not None:
            torch.cuda.current_stream(self.device).synchronize()  # the tables are read until here
            if d_spectra is not None:
                desc.spectra = None
            if d_weights is not None:
                desc.weights = None
                desc.weights_n = 0
        return d_cube


def _mix32(h):
    """murmur3's 32-bit finaliser on an int64 tensor of values in [0, 2^32):
    a bijection of 32-bit words, formed without leaving int64 range (each
    product is split into 16-bit halves of the constant)."""
    M = 0xFFFFFFFF

    def mul(x, c):
        return (x * (c & 0xFFFF) + (((x * (c >> 16)) & 0xFFFF) << 16)) & M

    h = h ^ (h >> 16)
    h = mul(h, 0x85EBCA6B)
    h = h ^ (h >> 13)
    h = mul(h, 0xC2B2AE35)
    return h ^ (h >> 16)


def record_permutation(n: int, device):
    """The fixed pseudo-random order of n records: the indices sorted by a
    bijective 32-bit hash of themselves.  The keys are distinct, so the order
    is the same on every device and on the CPU."""
    import torch

    k = torch.arange(n, dtype=torch.int64, device=device)
    return torch.argsort(_mix32(_mix32(k) ^ 0x6D617274))


def shuffle_records(d_rec):
    """Records (uint8 tensor of n * 48 bytes) in the fixed pseudo-random order
    of ``record_permutation``, gathered on the records' device.

    The FP32 gather's coarse sort keeps a cell's particles in record order
    within each bucket of its index hash, and takes them 32 at a time.
    Records that arrive spatially ordered (as ordered snapshots do) would give
    spatially compact chunks, which load one or two of the gather's four
    waves; shuffled records give chunks spread over the tile whatever the input
    order.  A structured reordering (a golden-ratio or other Weyl sweep) would
    interact with the multiplicative index hash.  The inserted cube depends on
    the order only through the FP32 summation order.

    The shuffle needs, for a moment, a second copy of the records and the
    permutation with its sort space (taken as 32 bytes per record); when the
    device has not that much memory free (with a 1 GiB margin), the records
    keep the given order."""
    import torch

    n = d_rec.numel() // _lib.PARTICLE_DTYPE.itemsize
    if n <= 1:
        return d_rec
    if d_rec.device.type == "cuda":
        free, _ = torch.cuda.mem_get_info(d_rec.device)
        free += torch.cuda.memory_reserved(d_rec.device) - torch.cuda.memory_allocated(d_rec.device)
        if free < n * (_lib.PARTICLE_DTYPE.itemsize + 32) + (1 << 30):
            return d_rec
    idx = record_permutation(n, d_rec.device)
    words = d_rec.view(torch.int32).view(n, _lib.PARTICLE_DTYPE.itemsize // 4)
    return words.index_select(0, idx).view(-1).view(torch.uint8)


def _spread4(v):
    """Bits 0..3 of v spread to the even bit positions 0, 2, 4, 6."""
    v = v & 15
    return (v & 1) | ((v & 2) << 1) | ((v & 4) << 2) | ((v & 8) << 3)


def cell_tile(nc):
    """The binning's pixel tile (rows along x, columns along y) for a cube of
    `nc` channels (mi_device.h wide_cells): 8 x 16 from 128 channels on, where
    cells take 128-channel blocks, else 16 x 16."""
    return (8, 16) if nc is not None and nc >= 128 else (16, 16)


def spatial_records(d_rec, nc=None):
    """Records (uint8 tensor of n * 48 bytes) sorted by the binning's pixel tile
    (``cell_tile(nc)``) of their pixel (ix, iy), tiles in its cell order
    (x-major), and by the pixel's Z order inside the tile (x and y bits
    interleaved, x on the even bits).  The MFMA form's binning keeps the
    record order within each cell, so its 32-particle chunks are then compact
    patches of the tile, and a chunk reads neighbouring records, whole cache
    lines instead of one 48-byte record per line.  Equal keys keep the given order (a stable sort),
    so the layout is the same on every device."""
    import torch

    n = d_rec.numel() // _lib.PARTICLE_DTYPE.itemsize
    if n <= 1:
        return d_rec
    words = d_rec.view(torch.int32).view(n, _lib.PARTICLE_DTYPE.itemsize // 4)
    ix = words[:, 0].to(torch.int64)
    iy = words[:, 1].to(torch.int64)
    off = 1 << 17  # |ix|, |iy| < 2^21: tile indices shifted to be non-negative
    tx, ty = cell_tile(nc)
    sx, sy = tx.bit_length() - 1, ty.bit_length() - 1
    key = ((((ix >> sx) + off) << 18) | ((iy >> sy) + off)) << 8
    key = key | _spread4(ix & (tx - 1)) | (_spread4(iy & (ty - 1)) << 1)
    idx = torch.argsort(key, stable=True)
    return words.index_select(0, idx).view(-1).view(torch.uint8)


def layout_records(d_rec, method="plain", nc=None):
    """The records' HBM layout for a gather method: shuffled for the FP32 form
    (``shuffle_records``: its chunks should spread over the tile), in cell and
    pixel Z order for the MFMA form (``spatial_records``: its chunks are
    compact, and neighbouring reads then share cache lines).  The inserted cube
    depends on the layout only through the FP32 summation order."""
    if _lib.METHODS[method] == _lib.METHOD_PLAIN_F32:
        return shuffle_records(d_rec)
    return spatial_records(d_rec, nc)


_default = {}


def default_inserter() -> Inserter:
    torch = _torch()
    dev = torch.cuda.current_device()
    if dev not in _default:
        _default[dev] = Inserter()
    return _default[dev]
