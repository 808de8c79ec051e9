"""Spectral models without a device form (MI_SPECTRUM_TABLE): a user subclass
of _BaseSpectrum supplies ``spectral_function`` (martini/spectral_models.py:
203-301); its spectra are evaluated on the host as the reference's
init_spectra does (:63-200) and the insertion kernels read them from HBM.

Pinned by (a) a Gaussian line written through spectral_function with scipy's
erf, against the oracle's FP64 Gaussian spectra and the oracle cube, and
(b) a top-hat line, whose channel fractions and total flux have closed
forms, against the oracle's pixel loop fed the same spectra.
"""

import numpy as np
import pytest
from scipy.special import erf

from martini_amd import DataCube, Martini, SPHSource, _lib, _pack, device_setup, synthetic
from martini_amd.sph_kernels import CubicSplineKernel
from martini_amd.spectral_models import K_B, M_P, GaussianSpectrum, _BaseSpectrum
from oracle import martini_oracle as O
from tests.helpers import FLUX_TOL, VOXEL_TOL, oracle_cube, parity


class ErfGaussian(_BaseSpectrum):
    """The thermal Gaussian line, as a user would write it."""

    def half_width(self, source):
        return np.sqrt(K_B * np.asarray(source.T_g, float) / M_P) / 1000.0

    def get_spectral_function_extra_data(self, source, datacube, mask=np.s_[...]):
        return {"sigma": np.asarray(self.half_width(source), float)[mask]}

    def spectral_function(self, a, b, vmids, extra_data=None):
        sig = extra_data["sigma"][:, np.newaxis]
        return 0.5 * (erf((b - vmids) / (np.sqrt(2.0) * sig))
                      - erf((a - vmids) / (np.sqrt(2.0) * sig)))


class TopHat(_BaseSpectrum):
    """A uniform line of half-width w km/s."""

    def __init__(self, w=6.0):
        self.w = float(w)
        super().__init__()

    def half_width(self, source):
        return self.w

    def spectral_function(self, a, b, vmids, extra_data=None):
        lo = np.maximum(a, vmids - self.w)
        hi = np.minimum(b, vmids + self.w)
        return np.clip(hi - lo, 0.0, None) / (2.0 * self.w)


def _top_hat_fractions(edges_kms, v, w):
    lo_e = np.minimum(edges_kms[:-1], edges_kms[1:])
    hi_e = np.maximum(edges_kms[:-1], edges_kms[1:])
    lo = np.maximum(lo_e[None, :], v[:, None] - w)
    hi = np.minimum(hi_e[None, :], v[:, None] + w)
    return np.clip(hi - lo, 0.0, None) / (2.0 * w)


def _martini(spectral_model, n=3000):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    src, dc, k, _ = synthetic.build(w, n)
    return w, Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=spectral_model,
                      quiet=True)


def test_erf_gaussian_host_spectra_match_oracle():
    w, m = _martini(ErfGaussian())
    m.init_spectra()
    P = O.prepare(*_arrays(w, 3000), distance_mpc=w.distance,
                  cube=O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                                  spectral_centre=w.vsys),
                  kernel=O.adaptive(O.KernelSpec("cubic")), rotation=w.rotation.as_matrix())
    ref = O.particle_spectra(P)
    assert m.spectral_model.spectra.shape == ref.shape
    scale = np.abs(ref).max()
    assert np.abs(m.spectral_model.spectra - ref).max() / scale < 1e-12


def _arrays(w, n):
    a = w.arrays(n)
    return a.xyz, a.vxyz, a.mHI, a.hsm, a.T


def test_top_hat_flux_closed_form():
    """Each line's channel fluxes add up to mHI / d^2 / 2.36e5 / |dv| times
    the channel width it covers: all of it for lines inside the band."""
    w, m = _martini(TopHat(6.0))
    S = m.spectral_model._host_spectra(m.source, m.datacube)
    dv = w.channel_width
    v = np.asarray(m.source.skycoords.radial_velocity, float)
    edges = m.datacube.velocity_channel_edges / 1e3
    inside = (v - 6.0 > edges.min()) & (v + 6.0 < edges.max())
    assert inside.sum() > 100
    expect = m.source.mHI_g / m.source.skycoords.distance**2 / 2.36e5
    np.testing.assert_allclose((S * dv).sum(axis=1)[inside], expect[inside], rtol=1e-12)
    # a line spans at most ceil(2 w / dv) + 1 channels
    assert ((S > 0).sum(axis=1) <= int(np.ceil(12.0 / dv)) + 1).all()


def test_table_records_and_descriptor():
    w, m = _martini(TopHat(6.0), n=500)
    rec, desc = _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, m.datacube)
    assert desc.spectrum == _lib.SPECTRUM_TABLE and desc.spectra_ld == w.n_channels
    S = desc.host_spectra
    assert S.dtype == np.float32 and S.shape == (m.source.npart, w.n_channels)
    np.testing.assert_array_equal(rec["k0"], np.arange(m.source.npart))
    assert (rec["amp"] == 1.0).all()
    for p in range(0, m.source.npart, 37):
        nz = np.flatnonzero(S[p])
        if nz.size:
            assert (rec["c_lo"][p], rec["c_hi"][p]) == (nz[0], nz[-1])
        else:
            assert rec["c_lo"][p] > rec["c_hi"][p]
    # a table descriptor without a table is an argument error
    bare = _lib.make_desc(desc.nx, desc.ny, desc.nc, _lib.SPECTRUM_TABLE)
    with pytest.raises(RuntimeError):
        _lib.count_pairs(rec, bare)


def test_model_without_spectral_function_raises():
    class HalfWidthOnly(_BaseSpectrum):
        def half_width(self, source):
            return 5.0

    w, m = _martini(HalfWidthOnly(), n=200)
    with pytest.raises(NotImplementedError):
        m.init_spectra()


def test_device_setup_rejects_table_spectra():
    dc = DataCube(n_px_x=16, n_px_y=16, n_channels=16, px_size=10.0, channel_width=4.0)
    with pytest.raises(NotImplementedError):
        device_setup.observation(distance=3.0, datacube=dc, sph_kernel=CubicSplineKernel(),
                                 spectral_model=TopHat())


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_erf_gaussian_table_matches_oracle(method):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    ref, _ = oracle_cube(w)
    src, dc, k, _ = synthetic.build(w)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=ErfGaussian(), quiet=True)
    m.insert_source_in_cube(method=method)
    voxel, flux = parity(m.datacube._array, ref)
    print(f"erf-gaussian table {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL and flux <= FLUX_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_top_hat_table_matches_oracle_pixel_loop(method):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    n, half = 2000, 6.0
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys)
    # sigma = half: the oracle's spectral pruning keeps lines within 4 half-widths
    P = O.prepare(*_arrays(w, n), distance_mpc=w.distance, cube=cube,
                  kernel=O.adaptive(O.KernelSpec("cubic")), sigma=half,
                  rotation=w.rotation.as_matrix())
    edges = cube.velocity_channel_edges() / 1e3
    width = np.abs(np.diff(edges))
    S = (P.mHI * P.distance_mpc**-2.0)[:, None] * _top_hat_fractions(edges, P.vlos_kms, half) \
        / width[None, :] / 2.36e5
    ref = O.cube_pixel_loop(P, S) / w.px_size**2
    src, dc, k, _ = synthetic.build(w, n)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=TopHat(half), quiet=True)
    assert m.source.npart == P.vlos_kms.size
    m.insert_source_in_cube(method=method)
    voxel, flux = parity(m.datacube._array, ref)
    print(f"top-hat table {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL and flux <= FLUX_TOL


@pytest.mark.gpu
def test_spectra_entry_point_reads_the_table():
    import torch

    w, m = _martini(TopHat(6.0), n=300)
    rec, desc = _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, m.datacube,
                                     out_scale=0.5)
    d_tab = torch.from_numpy(desc.host_spectra).cuda()
    desc.spectra = d_tab.data_ptr()
    got = _lib.spectra(rec, desc)
    np.testing.assert_array_equal(got, desc.host_spectra)
