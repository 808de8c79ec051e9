"""CPU tests of martini_amd's host side: FP64 setup against the oracle, record
packing against the reference's mask and window semantics, the C-ABI
library's exports and argument checks, and the reference's exceptions."""

import ctypes
import math
import re

import numpy as np
import pytest

from martini_amd import DataCube, L_coords, Martini, SPHSource, _lib, _pack, synthetic
from martini_amd import sph_kernels as K
from martini_amd.spectral_models import DiracDeltaSpectrum, GaussianSpectrum
from oracle import martini_oracle as O
from tests.conftest import ROOT


def _oracle_prepare(w, n=None, kernel=None):
    a = w.arrays(n)
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys)
    return O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                     kernel=kernel or O.adaptive(O.KernelSpec(w.kernel)),
                     rotation=w.rotation.as_matrix())


@pytest.mark.parametrize("name", ["disc1e4_64px", "hetero3e6_1024px"])
def test_setup_matches_oracle(name):
    w = synthetic.WORKLOADS[name]
    n = None if w.n <= 10_000 else 20_000
    src, dc, k, s = synthetic.build(w, n)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    P = _oracle_prepare(w, n)
    np.testing.assert_allclose(m.source.pixcoords, P.pixcoords, rtol=0, atol=1e-9)
    np.testing.assert_allclose(m.sph_kernel.sm_lengths, P.h_px, rtol=1e-14)
    assert np.array_equal(m.sph_kernel.sm_ranges, P.sm_ranges)
    assert np.array_equal(m.sph_kernel.kernel_indices, P.kernel_indices)
    np.testing.assert_allclose(m.source.skycoords.radial_velocity, P.vlos_kms, rtol=1e-13)


def test_setup_offset_pointing_and_padding():
    w = synthetic.WORKLOADS["disc1e4_64px"]
    a = w.arrays(3000)
    src = SPHSource(distance=5.0, rotation=w.rotation, ra=35.0, dec=-20.0, vpeculiar=30.0, T_g=a.T,
                    mHI_g=a.mHI, xyz_g=a.xyz, vxyz_g=a.vxyz, hsm_g=a.hsm)
    dc = DataCube(n_px_x=40, n_px_y=48, n_channels=24, px_size=20.0, channel_width=6.0,
                  spectral_centre=380.0, ra=35.01, dec=-19.98)
    dc.add_pad((3, 5))
    m = Martini(source=src, datacube=dc, sph_kernel=K.WendlandC2Kernel(),
                spectral_model=GaussianSpectrum(sigma="thermal"), quiet=True)
    cube = O.CubeSpec(40, 48, 24, 20.0, 6.0, spectral_centre=380.0, ra_deg=35.01, dec_deg=-19.98,
                      padx=3, pady=5)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=5.0, cube=cube,
                  kernel=O.adaptive(O.KernelSpec("wc2")), ra_deg=35.0, dec_deg=-20.0,
                  vpec_kms=30.0, rotation=w.rotation.as_matrix())
    np.testing.assert_allclose(m.source.pixcoords, P.pixcoords, atol=1e-9)
    assert np.array_equal(m.sph_kernel.sm_ranges, P.sm_ranges)


def test_frequency_mode_pixel_coordinates_match_oracle():
    """FREQ cubes: line-of-sight velocities become frequencies (radio
    convention) before the spectral pixel is formed."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    a = w.arrays(3000)
    src = SPHSource(distance=5.0, rotation=w.rotation, ra=35.0, dec=-20.0, vpeculiar=30.0, T_g=a.T,
                    mHI_g=a.mHI, xyz_g=a.xyz, vxyz_g=a.vxyz, hsm_g=a.hsm)
    dc = DataCube(n_px_x=40, n_px_y=48, n_channels=24, px_size=20.0, channel_width=2.5e4,
                  spectral_centre=380.0, ra=35.01, dec=-19.98, freq_channels=True)
    m = Martini(source=src, datacube=dc, sph_kernel=K.WendlandC2Kernel(),
                spectral_model=GaussianSpectrum(sigma="thermal"), quiet=True)
    cube = O.CubeSpec(40, 48, 24, 20.0, 2.5e4, spectral_centre=380.0, ra_deg=35.01, dec_deg=-19.98,
                      freq_mode=True)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=5.0, cube=cube,
                  kernel=O.adaptive(O.KernelSpec("wc2")), ra_deg=35.0, dec_deg=-20.0,
                  vpec_kms=30.0, rotation=w.rotation.as_matrix())
    assert m.source.npart == P.pixcoords.shape[1] > 0
    np.testing.assert_allclose(m.source.pixcoords, P.pixcoords, atol=1e-8)


def test_frequency_channels_match_oracle():
    dc = DataCube(n_px_x=8, n_px_y=8, n_channels=16, px_size=15.0, channel_width=18920.0,
                  spectral_centre=210.0, freq_channels=True)
    cube = O.CubeSpec(8, 8, 16, 15.0, 18920.0, spectral_centre=210.0, freq_mode=True)
    np.testing.assert_allclose(dc.velocity_channel_edges, cube.velocity_channel_edges(), rtol=1e-15)
    np.testing.assert_allclose(dc.channel_edges, cube.channel_edges(), rtol=1e-15)


def test_L_coords_matches_oracle():
    w = synthetic.WORKLOADS["disc1e4_64px"]
    a = w.arrays(2000)
    src = SPHSource(distance=3.0, L_coords=L_coords(incl=45.0, az_rot=30.0, pa=200.0), T_g=a.T,
                    mHI_g=a.mHI, xyz_g=a.xyz, vxyz_g=a.vxyz, hsm_g=a.hsm)
    rot = O.L_coords_matrix(a.xyz, a.vxyz, a.mHI, 45.0, 30.0, 200.0)
    np.testing.assert_allclose(src.current_rotation, rot, atol=1e-13)


def _records(w, n=None):
    src, dc, k, s = synthetic.build(w, n)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    rec, desc = _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, dc)
    return m, rec, desc


def test_record_footprints_reproduce_square_mask():
    """[ix - r + (fx > 0), ix + r] is exactly {i : |i - x| <= r}
    (martini/martini.py:271-274), clipped to the cube."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    m, rec, desc = _records(w)
    x, y = m.source.pixcoords[:2]
    r = m.sph_kernel.sm_ranges
    kinds = rec["r_kind"] & 15
    rr = rec["r_kind"] >> 8
    grid = np.arange(desc.nx)
    for p in np.flatnonzero(kinds != _lib.KERNEL_DIRAC)[:800]:
        ref = grid[np.abs(grid - x[p]) <= r[p]]
        lo = max(rec["ix"][p] - rr[p] + (rec["fx"][p] > 0), 0)
        hi = min(rec["ix"][p] + rr[p], desc.nx - 1)
        assert list(range(lo, hi + 1)) == list(ref)
        refy = np.arange(desc.ny)[np.abs(np.arange(desc.ny) - y[p]) <= r[p]]
        loy = max(rec["iy"][p] - rr[p] + (rec["fy"][p] > 0), 0)
        hiy = min(rec["iy"][p] + rr[p], desc.ny - 1)
        assert list(range(loy, hiy + 1)) == list(refy)


def test_record_dirac_pixel():
    """Dirac kernel: the one pixel with |dx| < 0.5 and |dy| < 0.5 (strict)."""
    dc = DataCube(n_px_x=8, n_px_y=8, n_channels=4, px_size=15.0, channel_width=4.0)
    xs = np.array([2.3, 2.5, 3.49999, 7.6, -0.3])
    src = SPHSource(distance=3.0, h=0.0, T_g=np.full(5, 1e4), mHI_g=np.full(5, 1e4),
                    xyz_g=np.zeros((5, 3)), vxyz_g=np.zeros((5, 3)), hsm_g=np.full(5, 0.01))
    src._init_skycoords()
    src.pixcoords = np.vstack((xs, np.full(5, 4.2), np.full(5, 1.0)))
    k = K.DiracDeltaKernel()
    k._init_sm_lengths(src, dc)
    k._init_sm_ranges()
    ix, iy, fx, fy, h, r_kind, _, _ = _pack.kernel_fields(k, src.pixcoords, 8, 8)
    assert list(ix[[0, 2, 4]]) == [2, 3, 0]
    assert list(r_kind[[0, 2, 4]] >> 8) == [0, 0, 0]
    assert (r_kind[1] >> 8) == -1  # exactly on a pixel edge: no pixel (strict <)
    assert (r_kind[3] >> 8) == -1  # nearest pixel 8 is outside the 8-pixel cube
    assert np.all(fx[[0, 2, 4]] == 0.0)


def test_record_channel_windows_cover_the_line():
    """Channels outside [c_lo, c_hi] carry at most erfc(4) (1.5e-8, both tails)
    of the particle's whole line (amp = mHI/d^2/dv/2.36e5 per unit profile) in
    the oracle's FP64 spectrum: the window reaches 4/s channel widths past the
    line centre on either side."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    m, rec, desc = _records(w)
    P = _oracle_prepare(w)
    S = O.particle_spectra(P)
    c = np.arange(desc.nc)[None, :]
    outside = (c < rec["c_lo"][:, None]) | (c > rec["c_hi"][:, None])
    amp = P.mHI / P.distance_mpc**2 / w.channel_width / O.MHI_TO_JY
    assert np.all((S * outside).sum(axis=1) <= math.erfc(4.0) * (1.0 + 1e-9) * amp)


def test_record_spectral_position_exact():
    """k0 + fu reproduces (v - e0)/de to FP32 precision of the fraction."""
    w = synthetic.WORKLOADS["disc1e4_64px"]
    m, rec, desc = _records(w)
    e, e0, de = _pack.channel_grid(m.datacube)
    u = (m.source.skycoords.radial_velocity * 1e3 - e0) / de
    np.testing.assert_allclose(rec["k0"] + rec["fu"].astype(float), u, rtol=0, atol=1e-7)


def test_dirac_spectrum_window_edge_inclusive():
    """A velocity exactly on an edge lies in both adjacent channels (the
    reference's heaviside(0) = 1, spectral_models.py:556-570)."""
    dc = DataCube(n_px_x=4, n_px_y=4, n_channels=8, px_size=10.0, channel_width=5.0,
                  spectral_centre=100.0)
    e = dc.velocity_channel_edges / 1e3
    k0, fu, s, amp, c_lo, c_hi = _pack.spectral_fields(
        np.array([e[3], 0.5 * (e[4] + e[5])]), 0.0, np.ones(2), np.ones(2), dc,
        _lib.SPECTRUM_DIRAC)
    assert (c_lo[0], c_hi[0]) == (2, 3)
    assert (c_lo[1], c_hi[1]) == (4, 4)


def _footprints(rec, nx, ny):
    """The device's clipped footprint (mi_device.h footprint) in numpy: the
    square mask of range r, cut for the tabulated kernels to the rows and
    columns within h (1 + 2^-12) of the particle, as offsets from ix (iy)
    formed in FP32 from the fraction fx (fy)."""
    rr = rec["r_kind"] >> 8
    ilo = rec["ix"] - rr + (rec["fx"] > 0)
    ihi = rec["ix"] + rr
    jlo = rec["iy"] - rr + (rec["fy"] > 0)
    jhi = rec["iy"] + rr
    h = rec["h"]
    cut = ((rec["r_kind"] & 15) <= _lib.KERNEL_QUARTIC) & (h > 0) & (h <= rr.astype(np.float32))
    R = h * np.float32(1.0 + 2.0**-12)  # FP32 throughout
    fx, fy = rec["fx"], rec["fy"]
    ix, iy = rec["ix"].astype(np.int64), rec["iy"].astype(np.int64)
    ilo = np.where(cut, np.maximum(ilo, ix + np.ceil(fx - R).astype(np.int64)), ilo)
    ihi = np.where(cut, np.minimum(ihi, ix + np.floor(fx + R).astype(np.int64)), ihi)
    jlo = np.where(cut, np.maximum(jlo, iy + np.ceil(fy - R).astype(np.int64)), jlo)
    jhi = np.where(cut, np.minimum(jhi, iy + np.floor(fy + R).astype(np.int64)), jhi)
    ilo, jlo = np.maximum(ilo, 0), np.maximum(jlo, 0)
    ihi, jhi = np.minimum(ihi, nx - 1), np.minimum(jhi, ny - 1)
    return rr, ilo, ihi, jlo, jhi, cut


@pytest.mark.parametrize("name", ["disc1e4_64px", "hetero3e6_1024px", "disc1e6_256px"])
def test_pair_count_matches_numpy(name):
    w = synthetic.WORKLOADS[name]
    m, rec, desc = _records(w, None if w.n <= 10_000 else 30_000)
    rr, ilo, ihi, jlo, jhi, cut = _footprints(rec, desc.nx, desc.ny)
    assert cut.mean() > 0.5  # the cut applies to most particles here
    clo, chi = np.maximum(rec["c_lo"], 0), np.minimum(rec["c_hi"], desc.nc - 1)
    ok = (rr >= 0) & (ilo <= ihi) & (jlo <= jhi) & (clo <= chi)
    # cells (mi_device.h): 8 x 16-pixel tiles x 128-channel blocks from 128
    # channels on, else 16 x 16 x 64
    tx, ty, cb = (8, 16, 128) if desc.nc >= 128 else (16, 16, 64)
    n = ((ihi // tx - ilo // tx + 1) * (jhi // ty - jlo // ty + 1) * (chi // cb - clo // cb + 1))
    assert _lib.count_pairs(rec, desc) == int(n[ok].sum())


@pytest.mark.parametrize("kernel", ["cubic", "wc2", "wc6", "quartic"])
def test_cut_footprint_drops_only_zero_weights(kernel):
    """Every pixel of the square mask that the device's footprint leaves out
    has weight exactly zero in the oracle's kernel integral, for h drawn up to
    the mask's range; and the footprint keeps pixels at q just below 1."""
    spec = O.KernelSpec(kernel)
    rng = np.random.default_rng(5)
    n = 4000
    rr = rng.integers(1, 9, n)
    h = (rr * rng.uniform(0.3, 1.0, n)).astype(np.float32)
    h[:8] = rr[:8]  # h = r exactly: the cut meets the mask
    x = rng.uniform(10, 20, n)
    y = rng.uniform(10, 20, n)
    rec = np.zeros(n, _lib.PARTICLE_DTYPE)
    rec["ix"], rec["iy"] = np.floor(x), np.floor(y)
    rec["fx"] = (x - np.floor(x)).astype(np.float32)
    rec["fy"] = (y - np.floor(y)).astype(np.float32)
    rec["h"] = h
    rec["r_kind"] = (rr << 8) | _lib.KERNEL_CUBIC
    x = rec["ix"] + rec["fx"].astype(float)
    y = rec["iy"] + rec["fy"].astype(float)
    _, ilo, ihi, jlo, jhi, cut = _footprints(rec, 64, 64)
    assert cut.all()
    dropped = kept_edge = 0
    for p in range(n):
        I = np.arange(np.ceil(x[p] - rr[p]), np.floor(x[p] + rr[p]) + 1)[:, None]
        J = np.arange(np.ceil(y[p] - rr[p]), np.floor(y[p] + rr[p]) + 1)[None, :]
        I, J = I + 0 * J, J + 0 * I  # the square mask, |i - x| <= r and |j - y| <= r
        kept = (I >= ilo[p]) & (I <= ihi[p]) & (J >= jlo[p]) & (J <= jhi[p])
        out = ~kept
        wts = spec.integral(x[p] - I[out], y[p] - J[out], float(h[p]))
        assert np.all(wts == 0.0), (p, wts.max())
        dropped += int(out.sum())
        q = np.hypot(x[p] - I, y[p] - J) / float(h[p])
        edge = (q < 1) & (q > 0.97)
        assert np.all(kept[edge])
        kept_edge += int(edge.sum())
    assert dropped > 10 * n and kept_edge > n


def test_library_exports_every_declared_symbol():
    header = open(f"{ROOT}/include/martini_mi355x.h").read()
    names = set(re.findall(r"\b(mi_[a-z_0-9]+)\s*\(", header))
    assert names >= {"mi_insert", "mi_insert_workspace_size", "mi_count_pairs_host",
                     "mi_kernel_weights", "mi_spectra", "mi_strerror", "mi_abi_version"}
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in sorted(names):
        assert hasattr(L, name), name
    assert _lib.lib().mi_abi_version() == _lib.ABI_VERSION == 3


def test_abi_argument_errors():
    L = _lib.lib()
    out = ctypes.c_size_t(0)
    assert L.mi_insert_workspace_size(None, 1, 1, ctypes.byref(out)) == -1
    d = _lib.make_desc(0, 4, 4, 0)
    assert L.mi_insert_workspace_size(ctypes.byref(d), 1, 1, ctypes.byref(out)) == -1
    d = _lib.make_desc(4, 4, 4, 7)  # unknown spectrum
    assert L.mi_count_pairs_host(None, 0, ctypes.byref(d)) == -1
    assert L.mi_kernel_weights(0, 0.0, 1.0, None, None, None, None, -1, None) == -1
    assert L.mi_kernel_weights(9, 0.0, 1.0, None, None, None, None, 0, None) == -4
    d = _lib.make_desc(4, 4, 4, 0)
    assert L.mi_insert(None, 0, 0, ctypes.byref(d), None, None, 0, 0, None) == -1
    # a workspace not 256-byte aligned is refused before any device work
    for entry in (L.mi_bin, L.mi_gather):
        args = (None, 0, 0, ctypes.byref(d)) + ((None,) if entry is L.mi_gather else ())
        assert entry(*args, ctypes.c_void_p(0x10010), ctypes.c_size_t(1 << 20), 0, None) == -1
    assert L.mi_strerror(-2) == b"workspace too small for this input"


def test_reference_exceptions():
    w = synthetic.WORKLOADS["disc1e4_64px"]
    a = w.arrays(200)
    with pytest.raises(RuntimeError, match="cannot guess coordinate_axis"):
        SPHSource(distance=3.0, mHI_g=np.ones(3), xyz_g=np.ones((3, 3)), vxyz_g=np.ones((3, 3)))
    with pytest.raises(ValueError, match="matching shapes"):
        SPHSource(distance=3.0, mHI_g=np.ones(4), xyz_g=np.ones((4, 3)), vxyz_g=np.ones((3, 4)))
    src = SPHSource(distance=3.0, T_g=a.T, mHI_g=a.mHI, xyz_g=a.xyz, vxyz_g=a.vxyz, hsm_g=a.hsm)
    with pytest.raises(ValueError, match="same length"):
        src.apply_mask(np.ones(3, bool))
    with pytest.raises(RuntimeError, match="No non-zero mHI"):
        src.apply_mask(np.zeros(200, bool))
    dc = DataCube(n_px_x=8, n_px_y=8, n_channels=4, px_size=15.0, channel_width=4.0)
    dc.add_pad((2, 2))
    with pytest.raises(RuntimeError, match="already padded"):
        dc.add_pad((1, 1))
    with pytest.raises(NotImplementedError):
        K.AdaptiveKernel()
    with pytest.raises(RuntimeError, match="truncation <2sigma"):
        K._GaussianKernel(truncate=1.0)


def test_validation_raises_before_launch():
    """A too-small smoothing length for a non-adaptive kernel raises the
    reference's RuntimeError (sph_kernels.py:188-189) on the host."""
    src = SPHSource(distance=3.0, T_g=np.full(1, 1e4), mHI_g=np.full(1, 1e4),
                    xyz_g=np.full((1, 3), 1e-3), vxyz_g=np.zeros((1, 3)), hsm_g=np.full(1, 0.9))
    dc = DataCube(n_px_x=100, n_px_y=100, n_channels=32, px_size=68.75493542, channel_width=10.0,
                  spectral_centre=src.vsys)
    m = Martini(source=src, datacube=dc, sph_kernel=K._CubicSplineKernel(),
                spectral_model=GaussianSpectrum(), quiet=True)
    with pytest.raises(RuntimeError, match="use this with care"):
        m.sph_kernel._confirm_validation(quiet=True)
    with pytest.raises(RuntimeError, match="use this with care"):
        m.insert_source_in_cube()


def test_product_kernel_constants_match_oracle():
    for cls, name in ((K._CubicSplineKernel, "cubic"), (K._WendlandC2Kernel, "wc2"),
                      (K._WendlandC6Kernel, "wc6"), (K._QuarticSplineKernel, "quartic")):
        assert cls()._rescale == pytest.approx(O.KernelSpec(name).rescale, rel=1e-14)
    g = K._GaussianKernel(truncate=6.0)
    assert g.norm == pytest.approx(O.gauss_norm(6.0), rel=1e-15)
    assert g.size_in_fwhm == pytest.approx(O.KernelSpec("gauss", truncate=6.0).size_in_fwhm,
                                           rel=1e-15)


def test_dirac_spectrum_pruning_and_records():
    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, _ = synthetic.build(w, 3000)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=DiracDeltaSpectrum(),
                quiet=True)
    rec, desc = _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, dc)
    width = rec["c_hi"] - rec["c_lo"] + 1
    assert np.all((width >= 0) & (width <= 2))
    assert desc.spectrum == _lib.SPECTRUM_DIRAC


def test_observation_struct_matches_host_objects():
    from martini_amd.device_setup import observation

    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, s = synthetic.build(w)
    o, desc = observation(distance=w.distance, datacube=dc, sph_kernel=k, spectral_model=s,
                          rotation=w.rotation)
    assert list(o.crpix) == list(dc.crpix) and list(o.cdelt) == list(dc.cdelt)
    e = dc.velocity_channel_edges
    assert o.e0_ms == e[0] and o.de_ms == (e[-1] - e[0]) / dc.n_channels
    assert np.allclose(np.array(o.rotation).reshape(3, 3), w.rotation.as_matrix(), atol=0)
    assert o.n_tiers == 3 and list(o.tier_kind) == [_lib.KERNEL_CUBIC, _lib.KERNEL_DIRAC,
                                                    _lib.KERNEL_GAUSS]
    for i, K in enumerate(k.kernels):
        assert o.tier_rescale[i] == K._rescale and o.tier_size[i] == K.size_in_fwhm
    assert o.tier_min[0] == 1.16 and o.tier_max[1] == 0.5 and o.tier_min[2] == 0.336
    assert o.sigma_kms < 0  # thermal
    assert (desc.nx, desc.ny, desc.nc) == (64, 64, 32)
    assert abs(desc.gauss_t[0] - 6.0) < 1e-6


def test_prepare_argument_errors():
    from martini_amd.device_setup import observation

    L = _lib.lib()
    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, s = synthetic.build(w)
    o, _ = observation(distance=w.distance, datacube=dc, sph_kernel=k, spectral_model=s)
    args = [None] * 6 + [0, None, None, None, None, 0, None]
    assert L.mi_prepare(*args) == -1  # no observation
    args[7] = ctypes.byref(o)
    assert L.mi_prepare(*args) == -1  # no counts buffer
    out = ctypes.c_size_t(0)
    assert L.mi_prepare_workspace_size(-1, ctypes.byref(out)) == -1
    o.n_tiers = 0
    args[9] = ctypes.c_void_p(8)
    assert L.mi_prepare(*args) == -1
    o.n_tiers, o.spectrum = 3, 7
    assert L.mi_prepare(*args) == -1


def test_global_profile_setup():
    from martini_amd import GlobalProfile
    from martini_amd.sph_kernels import DiracDeltaKernel

    w = synthetic.WORKLOADS["disc1e4_64px"]
    a = w.arrays(2000)
    src = SPHSource(distance=w.distance, rotation=w.rotation, T_g=a.T, mHI_g=a.mHI, xyz_g=a.xyz,
                    vxyz_g=a.vxyz, hsm_g=a.hsm)
    gp = GlobalProfile(source=src, spectral_model=GaussianSpectrum(sigma="thermal"),
                       n_channels=16, channel_width=4.0, spectral_centre=w.vsys, quiet=True)
    assert gp.datacube._array.shape == (1, 1, 16) and gp.datacube.px_size == 3600.0
    assert isinstance(gp.sph_kernel, DiracDeltaKernel)
    assert np.all(np.isinf(gp.sph_kernel.sm_ranges))
    assert np.all(gp.source.pixcoords[:2] == 0)
    # spectral pruning only: a narrow band drops the particles outside it
    assert 0 < gp.source.npart < 2000
    rec, desc = _pack.pack_particles(gp.source, gp.sph_kernel, gp.spectral_model, gp.datacube)
    assert np.all(rec["ix"] == 0) and np.all(rec["iy"] == 0) and np.all(rec["r_kind"] >> 8 == 0)
