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

    def upload(self, records: np.ndarray):
        """Copy host records into HBM (uint8 tensor of n * 48 bytes), in the
        order of ``shuffle_records``."""
        rec = np.ascontiguousarray(records)
        return shuffle_records(self.torch.from_numpy(rec.view(np.uint8)).to(self.device))

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
        n_pairs = _lib.count_pairs(records, desc)
        d_rec = self.upload(records)
        if d_cube is None:
            d_cube = torch.zeros(desc.nx * desc.ny * desc.nc, dtype=torch.float32, device=self.device)
        self.insert_device(d_rec, n, n_pairs, desc, d_cube, method=method)
        if d_spectra is not None:
            torch.cuda.current_stream(self.device).synchronize()  # the table is read until here
            desc.spectra = None
        return d_cube


SHUFFLE_SEED = 0x6D617274


def shuffle_records(d_rec):
    """Records (uint8 tensor of n * 48 bytes) in a fixed pseudo-random order,
    drawn and gathered on the records' device (``torch.randperm`` with a
    generator seeded by SHUFFLE_SEED, so the same n gives the same order).

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
    gen = torch.Generator(device=d_rec.device)
    gen.manual_seed(SHUFFLE_SEED)
    idx = torch.randperm(n, generator=gen, device=d_rec.device)
    words = d_rec.view(torch.int32).view(n, _lib.PARTICLE_DTYPE.itemsize // 4)
    return words.index_select(0, idx).view(-1).view(torch.uint8)


_default = {}


def default_inserter() -> Inserter:
    torch = _torch()
    dev = torch.cuda.current_device()
    if dev not in _default:
        _default[dev] = Inserter()
    return _default[dev]
