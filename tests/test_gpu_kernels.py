"""GPU unit tests of the device kernel integrals and spectra against the FP64
oracle, through the C ABI (mi_kernel_weights, mi_spectra)."""

import numpy as np
import pytest

from martini_amd import Martini, _lib, _pack, synthetic
from martini_amd import sph_kernels as K
from oracle import martini_oracle as O

pytestmark = pytest.mark.gpu

CASES = [(_lib.KERNEL_CUBIC, "cubic", 0.0), (_lib.KERNEL_WC2, "wc2", 0.0),
         (_lib.KERNEL_WC6, "wc6", 0.0), (_lib.KERNEL_QUARTIC, "quartic", 0.0),
         (_lib.KERNEL_GAUSS, "gauss", 3.0), (_lib.KERNEL_GAUSS, "gauss", 6.0)]


@pytest.mark.parametrize("code,name,t", CASES)
def test_kernel_weights_match_oracle(code, name, t):
    """Device weights against the FP64 oracle, relative to each particle's peak
    weight: 5e-7 for the tabulated closed forms (FP32 q and interpolation sum)
    and the Gaussian (a product of three FP32 erf factors); 5e-6 for
    WendlandC6, whose rounded reference coefficients leave a square-root
    edge at q -> 1 that the interpolation table smooths."""
    rng = np.random.default_rng(3)
    n = 200_000
    h = np.exp(rng.uniform(np.log(0.3), np.log(60.0), n)).astype(np.float32)
    dx = (rng.uniform(-1.1, 1.1, n) * h * 2.6).astype(np.float32)
    dy = (rng.uniform(-1.1, 1.1, n) * h * 2.6).astype(np.float32)
    spec = O.KernelSpec(name, truncate=t or None)
    norm = spec.norm if name == "gauss" else 1.0
    got = _lib.kernel_weights(code, dx, dy, h, t, norm)
    ref = spec.integral(dx.astype(float), dy.astype(float), h.astype(float))
    peak = spec.integral(np.zeros(n), np.zeros(n), h.astype(float))
    err = np.abs(got - ref) / peak
    assert err.max() <= (5e-6 if name == "wc6" else 5e-7), (name, err.max())


def test_dirac_weights():
    dx = np.array([0.0, 0.49, 0.5, -0.5, 0.3], np.float32)
    dy = np.array([0.0, -0.49, 0.0, 0.0, 0.6], np.float32)
    got = _lib.kernel_weights(_lib.KERNEL_DIRAC, dx, dy, np.ones(5, np.float32))
    assert list(got) == [1.0, 1.0, 0.0, 0.0, 0.0]


def test_px_weight_plugin_total():
    """tests/test_sph_kernels.py:44-72 through the product's _px_weight."""
    for cls in (K._CubicSplineKernel, K._WendlandC2Kernel, K._QuarticSplineKernel):
        k = cls()
        r = np.arange(0, 50)
        xg, yg = np.meshgrid(np.r_[r[1:][::-1], r], np.r_[r[1:][::-1], r])
        dij = np.vstack((xg.ravel(), yg.ravel())).astype(float)
        k.sm_lengths = np.full(dij.shape[1], 20.0)
        assert k._px_weight(dij).sum() == pytest.approx(1.0, rel=1e-3)


@pytest.mark.parametrize("spectrum", ["gauss", "dirac"])
def test_spectra_match_oracle(spectrum):
    from martini_amd.spectral_models import DiracDeltaSpectrum, GaussianSpectrum

    w = synthetic.WORKLOADS["disc1e4_64px"]
    sm = GaussianSpectrum(sigma="thermal") if spectrum == "gauss" else DiracDeltaSpectrum()
    src, dc, k, _ = synthetic.build(w, 3000)
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=sm, quiet=True)
    m.init_spectra()
    a = w.arrays(3000)
    cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                      spectral_centre=w.vsys)
    P = O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                  kernel=O.adaptive(O.KernelSpec("cubic")), rotation=w.rotation.as_matrix(),
                  spectrum=spectrum)
    ref = O.particle_spectra(P)
    amp = (P.mHI / P.distance_mpc**2 / w.channel_width / O.MHI_TO_JY)[:, None]
    err = np.abs(m.spectral_model.spectra - ref) / amp
    assert err.max() <= (1e-6 if spectrum == "gauss" else 1e-6)
