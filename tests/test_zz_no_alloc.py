"""The no-allocation entry point mi_insert_noalloc, called through the C ABI:
its cube equals mi_insert's bit for bit (same kernels, same workspace layout,
the static scratch in place of the caller's workspace), on two calls in a row,
the second with a smaller input, for both forms; an input whose workspace
exceeds the scratch is refused before any launch."""

import numpy as np
import pytest

from martini_amd import Martini, _lib, _pack, synthetic

pytestmark = pytest.mark.gpu


def _records(n):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, s = synthetic.build(w, n)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    return _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, dc,
                                out_scale=1.0 / dc.px_size**2)


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_noalloc_equals_insert(method):
    import torch

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    for n in (10_000, 2_500):  # the second call has a smaller input
        rec, desc = _records(n)
        n_rec = int(rec.size)
        n_pairs = _lib.count_pairs(rec, desc)
        assert _lib.workspace_bytes(desc, n_rec, n_pairs) <= _lib.SCRATCH_BYTES
        d_rec = torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8)).to(dev)
        size = desc.nx * desc.ny * desc.nc
        want = torch.zeros(size, dtype=torch.float32, device=dev)
        ws = torch.empty(_lib.workspace_bytes(desc, n_rec, n_pairs), dtype=torch.uint8,
                         device=dev)
        _lib.insert(d_rec, n_rec, n_pairs, desc, want, ws, method=method, stream=stream)
        got = torch.zeros(size, dtype=torch.float32, device=dev)
        rc = _lib.insert_noalloc(d_rec, n_rec, n_pairs, desc, got, method=method, stream=stream)
        assert rc == _lib.MI_OK, rc
        torch.cuda.synchronize()
        a, b = want.cpu().numpy(), got.cpu().numpy()
        print(f"{method} n={n_rec} pairs={n_pairs}: peak {a.max():.3e}, "
              f"max |noalloc - insert| {np.abs(a - b).max():.3e}", flush=True)
        assert a.max() > 0
        assert np.array_equal(a, b)


def test_noalloc_refuses_inputs_beyond_the_scratch():
    import torch

    rec, desc = _records(2_000)
    dev = torch.device("cuda", 0)
    d_rec = torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8)).to(dev)
    cube = torch.zeros(desc.nx * desc.ny * desc.nc, dtype=torch.float32, device=dev)
    huge = 200_000_000  # pair capacity whose workspace is far beyond 256 MiB
    assert _lib.workspace_bytes(desc, int(rec.size), huge) > _lib.SCRATCH_BYTES
    rc = _lib.insert_noalloc(d_rec, int(rec.size), huge, desc, cube, method="plain")
    assert rc == _lib.MI_ERR_SCRATCH
    torch.cuda.synchronize()
    assert float(cube.abs().max()) == 0.0  # nothing launched
