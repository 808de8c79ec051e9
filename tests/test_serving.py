"""Serving form: requests served concurrently give the same cubes as alone,
and a served request matches the FP64 oracle."""

import numpy as np
import pytest

from martini_amd import serving
from oracle import cube_c
from oracle import martini_oracle as O
from tests.helpers import FLUX_TOL, VOXEL_TOL, parity

pytestmark = pytest.mark.gpu


def _stream(n_requests=24):
    return serving.request_stream(n_requests, seed=5000, n_range=(1e4, 2e5), rate_hz=2000.0,
                                  shapes=serving.SHAPES[:3])


def test_mix_equals_alone():
    specs = _stream()
    sources = [serving.disc_on_device(s) for s in specs]
    mix = serving.Service(slots=4, method="mfma").run(specs, sources)
    summary = serving.summarize(mix)
    print("mix:", {k: round(v, 3) if isinstance(v, float) else v for k, v in summary.items()})
    alone_service = serving.Service(slots=1, method="mfma", fuse_max_n=0)
    overlapped = 0
    for r in mix:
        others = [q for q in mix if q.rid != r.rid and q.inserted < r.done and q.done > r.inserted]
        overlapped += bool(others)
    fused = 0
    for spec, src, r in zip(specs, sources, mix):
        spec0 = serving.RequestSpec(**{**spec.__dict__, "arrival": 0.0})
        (alone,) = alone_service.run([spec0], [src])
        assert alone.kept == r.kept and alone.group == 1
        if r.group == 1:
            assert alone.pairs == r.pairs
            assert np.array_equal(alone.cube, r.cube), spec
        else:  # fused: same particles, chunked with other requests' particles
            fused += 1
            voxel, flux = parity(r.cube.astype(float), alone.cube.astype(float))
            assert voxel <= VOXEL_TOL and flux <= FLUX_TOL, (spec, voxel, flux)
    print(f"{overlapped} of {len(mix)} requests overlapped another on the GPU, {fused} fused")
    assert overlapped > 0 and fused > 0


def test_served_cube_matches_oracle():
    spec = serving.RequestSpec(rid=0, n=20_000, seed=5100, incl=35.0, distance=7.0, n_px=64,
                               n_channels=32, kernel="wc2")
    src = serving.disc_on_device(spec)
    (res,) = serving.Service(slots=1, method="plain").run([spec], [src])
    a = {k: getattr(src, k).cpu().numpy() for k in ("xyz", "vxyz", "mHI", "hsm", "T")}
    cube = O.CubeSpec(spec.n_px, spec.n_px, spec.n_channels, spec.px_size, spec.channel_width,
                      spectral_centre=70.0 * spec.distance)
    from scipy.spatial.transform import Rotation

    P = O.prepare(a["xyz"], a["vxyz"], a["mHI"], a["hsm"], a["T"], distance_mpc=spec.distance,
                  cube=cube, kernel=O.adaptive(O.KernelSpec("wc2")),
                  rotation=Rotation.from_euler("y", np.deg2rad(spec.incl)).as_matrix())
    ref = cube_c.cube(P) / spec.px_size**2
    voxel, flux = parity(res.cube.astype(float), ref)
    print(f"served request vs oracle: voxel {voxel:.3e} flux {flux:.3e}, kept {res.kept}")
    assert res.kept == P.keep.sum()
    assert voxel <= VOXEL_TOL and flux <= FLUX_TOL
