"""SPH kernels without a device form (MI_KERNEL_TABLE): a user subclass of
_BaseSPHKernel, or a subclass that overrides _kernel_integral, runs through
the GPU insertion with its own _kernel_integral evaluated on the host per
masked (particle, pixel) pair, as the reference's pixel loop evaluates it
(martini/martini.py:243-283, _px_weight sph_kernels.py:85-119).

CPU tests pin the host side (kind selection, the weight table against a direct
per-pixel evaluation, the record offsets); GPU tests compare inserted cubes
with the FP64 oracle."""

import numpy as np
import pytest

from martini_amd import Martini, _lib, _pack, synthetic
from martini_amd import sph_kernels as K
from oracle import cube_c
from oracle import martini_oracle as O
from tests.helpers import FLUX_TOL, VOXEL_TOL, parity


def box_integral(dx, dy, h):
    """Pixel integral of a uniform square of side h centred on the particle:
    the overlap of [d - h/2, d + h/2] with the pixel [-1/2, 1/2] on each axis,
    over h^2.  Not a function of the radius alone."""
    def ov(d):
        return np.clip(np.minimum(d + h / 2, 0.5) - np.maximum(d - h / 2, -0.5), 0.0, None)
    return ov(np.asarray(dx, float)) * ov(np.asarray(dy, float)) / (h * h)


class BoxKernel(K._BaseSPHKernel):
    """A user kernel written against the reference's plugin contract."""

    min_valid_size = 1.0

    def __init__(self):
        super().__init__()
        self.size_in_fwhm = 1.0  # mask half-width ceil(h) covers h/2 + 1/2 for h >= 1

    def kernel(self, q):
        return np.where(np.abs(q) <= 0.5, 1.0, 0.0)

    def _kernel_integral(self, dij, h, mask=Ellipsis):
        dij = np.asarray(dij, float)
        return box_integral(dij[0], dij[1], np.asarray(h, float))

    def _validate(self, sm_lengths, noraise=False, quiet=False):
        valid = np.asarray(sm_lengths) >= self.min_valid_size
        if not valid.all() and not noraise:
            raise RuntimeError("BoxKernel: smoothing lengths below 1 px")
        return valid


class UserWendlandC2(K._WendlandC2Kernel):
    """A subclass that brings its own _kernel_integral: the WendlandC2 closed
    form written out in numpy, as a plugin author would."""

    def _kernel_integral(self, dij, h, mask=Ellipsis):
        dij = np.asarray(dij, float)
        h = np.asarray(h, float)
        R2 = (dij**2).sum(axis=0) / h**2
        out = np.zeros(R2.shape)
        use = (R2 < 1) & (R2 > 0)
        r2 = R2[use]
        A = np.sqrt(1 - r2)
        out[use] = (5 * r2**2 * (0.5 * r2 + 3) * np.log((1 + A) / np.sqrt(r2))
                    + A * (-13.5 * r2**2 - 14 / 3 * r2 + 2 / 3))
        out[R2 == 0] = 2 / 3
        return out * 21 / (2 * np.pi) / h**2


def user_wc2_adaptive():
    return K._AdaptiveKernel((UserWendlandC2(), K.DiracDeltaKernel(),
                              K._GaussianKernel(truncate=6.0)))


class _BoxSpec:
    """The box kernel as an oracle kernel spec (non-adaptive)."""

    name = "box"
    rescale = 1.0
    size_in_fwhm = 1.0

    def integral(self, dx, dy, h):
        return box_integral(dx, dy, h)


def test_device_kind_selection():
    assert _pack.device_kind(K._CubicSplineKernel()) == _lib.KERNEL_CUBIC
    assert _pack.device_kind(K._GaussianKernel(truncate=3.0)) == _lib.KERNEL_GAUSS
    assert _pack.device_kind(K.DiracDeltaKernel()) == _lib.KERNEL_DIRAC
    assert _pack.device_kind(BoxKernel()) == _lib.KERNEL_TABLE
    assert _pack.device_kind(UserWendlandC2()) == _lib.KERNEL_TABLE


def _martini(w, n, kernel):
    src, dc, _, s = synthetic.build(w, n, kernel=kernel)
    return Martini(source=src, datacube=dc, sph_kernel=kernel, spectral_model=s, quiet=True)


@pytest.mark.parametrize("make", [BoxKernel, user_wc2_adaptive], ids=["box", "wc2-subclass"])
def test_table_weights_match_per_pixel_evaluation(make):
    """Each table particle's weights are the kernel's own _kernel_integral at
    every pixel its square mask admits (|i - x| <= r, |j - y| <= r), row-major
    over the footprint clipped to the cube, starting at the offset its record
    carries in h."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    m = _martini(w, 3000, make())
    rec, desc = _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, m.datacube)
    table = desc.host_weights
    assert table.dtype == np.float32 and table.size > 0
    nx, ny = m.datacube._array.shape[:2]
    x, y = m.source.pixcoords[:2]
    r = np.asarray(m.sph_kernel.sm_ranges)
    h_eff = np.asarray(m.sph_kernel.sm_lengths) * np.asarray(m.sph_kernel._rescale)
    kinds = rec["r_kind"] & 15
    offs = rec.view(np.uint32).reshape(rec.size, 12)[:, 4]
    rng = np.random.default_rng(0)
    picked = rng.choice(np.flatnonzero(kinds == _lib.KERNEL_TABLE), 40, replace=False)
    kernels, which = m.sph_kernel._device_kernels()
    for p in picked:
        ii = np.arange(nx)[np.abs(np.arange(nx) - x[p]) <= r[p]]
        jj = np.arange(ny)[np.abs(np.arange(ny) - y[p]) <= r[p]]
        I, J = np.meshgrid(ii, jj, indexing="ij")
        dij = np.vstack(((x[p] - I).ravel(), (y[p] - J).ravel()))
        want = kernels[which[p]]._kernel_integral(dij, np.full(I.size, h_eff[p]))
        got = table[offs[p]: offs[p] + I.size]
        np.testing.assert_allclose(got, want.astype(np.float32), rtol=0, atol=0)
    # particles of built-in kinds keep their kernel size in h
    other = kinds != _lib.KERNEL_TABLE
    np.testing.assert_allclose(rec["h"][other], h_eff[other].astype(np.float32))


def test_box_kernel_conserves_weight():
    """The box kernel's weights over a footprint inside the cube sum to 1:
    the host table keeps every masked pixel."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    m = _martini(w, 2000, BoxKernel())
    rec, desc = _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, m.datacube)
    offs = rec.view(np.uint32).reshape(rec.size, 12)[:, 4].astype(np.int64)
    nx, ny = m.datacube._array.shape[:2]
    x, y = m.source.pixcoords[:2]
    r = np.asarray(m.sph_kernel.sm_ranges)
    inside = (x - r >= 0) & (y - r >= 0) & (x + r <= nx - 1) & (y + r <= ny - 1)
    ends = np.append(offs[1:], desc.host_weights.size)
    sums = np.add.reduceat(desc.host_weights.astype(np.float64), offs)
    assert np.all(ends > offs)
    np.testing.assert_allclose(sums[inside], 1.0, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_user_wc2_subclass_matches_oracle(method):
    """The WendlandC2 tier of the adaptive kernel replaced by a user subclass
    with its own integral: the cube equals the oracle's built-in WendlandC2
    cube within the usual bars (the 1e6-particle disc's geometry, 1e5
    particles: all three tiers occur)."""
    w = synthetic.WORKLOADS["disc1e6_256px"]
    n = 100_000
    m = _martini(w, n, user_wc2_adaptive())
    m.insert_source_in_cube(method=method)
    a = w.arrays(n)
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                  kernel=O.adaptive(O.KernelSpec("wc2")), rotation=w.rotation.as_matrix())
    ref = cube_c.cube(P) / w.px_size**2
    voxel, flux = parity(m.datacube._array, ref)
    tiers = np.bincount(np.asarray(m.sph_kernel.kernel_indices) + 1, minlength=4)
    print(f"user WC2 subclass {method}: voxel {voxel:.3e} flux {flux:.3e}, tiers {tiers}")
    assert voxel <= VOXEL_TOL and flux <= FLUX_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_box_kernel_matches_oracle_pixel_loop(method):
    """A non-radial user kernel against the oracle's pixel loop fed the same
    kernel integral (the reference's _evaluate_pixel_spectrum semantics)."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    n = 3000
    m = _martini(w, n, BoxKernel())
    m.insert_source_in_cube(method=method, skip_validation=True)  # the box is exact below 1 px
    a = w.arrays(n)
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                  kernel=_BoxSpec(), rotation=w.rotation.as_matrix())
    ref = O.cube_pixel_loop(P) / w.px_size**2
    voxel, flux = parity(m.datacube._array, ref)
    print(f"box kernel {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL and flux <= FLUX_TOL
