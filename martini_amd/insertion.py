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
        """Copy host records into HBM (uint8 tensor of n * 48 bytes)."""
        rec = np.ascontiguousarray(records)
        return self.torch.from_numpy(rec.view(np.uint8)).to(self.device)

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


_default = {}


def default_inserter() -> Inserter:
    torch = _torch()
    dev = torch.cuda.current_device()
    if dev not in _default:
        _default[dev] = Inserter()
    return _default[dev]
