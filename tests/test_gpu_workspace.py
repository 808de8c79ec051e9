"""Insertion does not depend on what the workspace held before the call.

Split cells leave partial tiles in the workspace, and the MFMA form writes
only the 1 KiB sub-tiles its work item reached (a per-tile mask says which);
the reduction must read nothing else.  A workspace filled with NaN bytes, and
the same workspace reused by a second, smaller input, must give the cube that
a zeroed workspace gives, bit for bit, for both forms."""

import numpy as np
import pytest

from martini_amd import Martini, _lib, _pack, synthetic

pytestmark = pytest.mark.gpu


def _records(workload, n):
    w = synthetic.WORKLOADS[workload]
    src, dc, k, s = synthetic.build(w, n)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    return _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, dc,
                                out_scale=1.0 / dc.px_size**2)


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_stale_workspace_is_never_read(method):
    import torch

    from martini_amd.insertion import layout_records

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    cases = [_records("disc1e6_256px", 200_000), _records("disc1e6_256px", 60_000)]
    need = max(_lib.workspace_bytes(d, int(r.size), _lib.count_pairs(r, d)) for r, d in cases)
    dirty = torch.full((need,), 0xFF, dtype=torch.uint8, device=dev)  # NaN in every float
    for rec, desc in cases:  # the second input reuses what the first left behind
        n_rec = int(rec.size)
        n_pairs = _lib.count_pairs(rec, desc)
        d_rec = torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8)).to(dev)
        d_rec = layout_records(d_rec, method, desc.nc)
        size = desc.nx * desc.ny * desc.nc
        clean = torch.zeros(_lib.workspace_bytes(desc, n_rec, n_pairs), dtype=torch.uint8,
                            device=dev)
        want = torch.zeros(size, dtype=torch.float32, device=dev)
        got = torch.zeros(size, dtype=torch.float32, device=dev)
        _lib.insert(d_rec, n_rec, n_pairs, desc, want, clean, method=method, stream=stream)
        _lib.insert(d_rec, n_rec, n_pairs, desc, got, dirty, method=method, stream=stream)
        split = _lib.status(desc, n_rec, n_pairs, dirty)[3]
        torch.cuda.synchronize()
        a, b = want.cpu().numpy(), got.cpu().numpy()
        print(f"{method} n={n_rec} pairs={n_pairs} split items {split}: "
              f"peak {a.max():.3e}, NaN in dirty-workspace cube {int(np.isnan(b).sum())}",
              flush=True)
        assert split > 0  # the case has split cells
        assert a.max() > 0
        assert np.array_equal(a, b)
