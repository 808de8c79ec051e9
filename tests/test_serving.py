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


def _check_mix_against_alone(specs, mix_service):
    """Serve specs as a mix, then each request alone: unfused requests give
    bitwise the cube they get alone, fused ones agree within the parity bars.
    Stops at the first mismatch.  Returns (mix results, overlapped, fused)."""
    sources = [serving.disc_on_device(s) for s in specs]
    mix = mix_service.run(specs, sources)
    alone_service = serving.Service(slots=1, method=mix_service.method, fuse_max_n=0)
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
    return mix, overlapped, fused


def test_mix_equals_alone_at_the_extremes():
    """The mix's extremes: 512^2 x 256 cubes and requests of 1e6 to 1e7
    particles, served concurrently with small fused requests (64^2 x 32 and
    128^2 x 64, 1e4 to 5e4 particles)."""
    big = [(1_000_000, 512, 256, "cubic"), (10_000_000, 512, 256, "wc2"),
           (3_000_000, 256, 128, "cubic"), (2_000_000, 512, 256, "cubic")]
    small = [(10_000, 64, 32, "cubic"), (30_000, 64, 32, "cubic"), (20_000, 64, 32, "cubic"),
             (50_000, 128, 64, "wc2"), (15_000, 128, 64, "wc2"), (12_000, 512, 256, "cubic")]
    rng = np.random.default_rng(77)
    specs = []
    for i, (n, n_px, nc, kern) in enumerate(big + small):
        specs.append(serving.RequestSpec(rid=i, n=n, seed=6000 + i,
                                         incl=float(rng.uniform(0.0, 90.0)),
                                         distance=float(rng.uniform(3.0, 30.0)), n_px=n_px,
                                         n_channels=nc, kernel=kern, arrival=0.0005 * i))
    mix, overlapped, fused = _check_mix_against_alone(specs, serving.Service(slots=4,
                                                                           method="mfma"))
    summary = serving.summarize(mix)
    print("extremes mix:", {k: round(v, 3) if isinstance(v, float) else v
                            for k, v in summary.items()})
    print(f"{overlapped} of {len(mix)} requests overlapped another on the GPU, {fused} fused")
    assert overlapped > 0 and fused > 0


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
