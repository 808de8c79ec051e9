"""Pins the oracle's projection, channels, spectra, pruning and cube.

Each test restates a known-answer test of the reference (file:line cited) or
checks a property the mathematics fixes.
"""

import numpy as np
import pytest
from scipy.special import erf

from oracle import cube_c
from oracle import martini_oracle as O

KPC_PER_ARCSEC_DIST = 1.0 / (1.0 / O.ARCSEC_PER_RAD) / 1000.0  # Mpc at which 1 kpc = 1 arcsec


def ref_matrices(ra_deg, dec_deg):
    """Matrices of tests/test_sources.py:245-262 (row-vector convention)."""
    ra, dec = np.deg2rad(ra_deg), np.deg2rad(dec_deg)
    R_y = np.array([[np.cos(dec), 0, np.sin(dec)], [0, 1, 0], [-np.sin(dec), 0, np.cos(dec)]])
    R_z = np.array([[np.cos(-ra), -np.sin(-ra), 0], [np.sin(-ra), np.cos(-ra), 0], [0, 0, 1]])
    return R_y.dot(R_z)


def test_init_pixcoords_lineup():
    """tests/test_sources.py:335-373: 1 kpc = 1 arcsec, h = 0."""
    xyz = np.vstack((np.zeros(6), np.linspace(-2.5, 2.5, 6), np.linspace(-2.5, 2.5, 6))).T
    vxyz = np.vstack((np.linspace(-2.5, 2.5, 6), np.zeros(6), np.zeros(6))).T
    sky = O.sky_coordinates(xyz, vxyz, KPC_PER_ARCSEC_DIST, 0.0, 0.0, h=0.0)
    cube = O.CubeSpec(6, 6, 6, 1.0, 1.0)
    pix = O.world_to_pixel(sky["ra_deg"], sky["dec_deg"], sky["vlos_kms"], cube)
    expected = np.vstack((np.arange(6)[::-1], np.arange(6), np.arange(6)[::-1]))
    np.testing.assert_allclose(pix, expected, atol=1e-4)


@pytest.mark.parametrize("ra", [0.0, 30.0, -30.0])
@pytest.mark.parametrize("dec", [0.0, 30.0, -30.0])
def test_ra_dec_rotation(ra, dec):
    """tests/test_sources.py:229-267: rotation before translation."""
    xyz = np.arange(3).reshape((1, 3)).astype(float)
    sky = O.sky_coordinates(xyz, xyz, 0.0, ra, dec, h=0.0)
    rot = ref_matrices(ra, dec)
    np.testing.assert_allclose(sky["xyz_mpc"] * 1000.0, xyz.dot(rot), atol=1e-12)
    np.testing.assert_allclose(sky["vxyz_kms"], xyz.dot(rot), atol=1e-12)


@pytest.mark.parametrize("ra", [0.0, 30.0, -30.0])
@pytest.mark.parametrize("dec", [0.0, 30.0, -30.0])
@pytest.mark.parametrize("distance", [0.0, 3.0])
@pytest.mark.parametrize("vpec", [0.0, 100.0])
def test_dist_vpec_translation(ra, dec, distance, vpec):
    """tests/test_sources.py:269-316."""
    xyz = np.arange(1, 4).reshape((1, 3)).astype(float)
    sky = O.sky_coordinates(xyz, xyz, distance, ra, dec, h=0.7, vpec_kms=vpec)
    rot = ref_matrices(ra, dec)
    direction = rot.T.dot(np.array([[1], [0], [0]]))
    pos = xyz.dot(rot) / 1000.0 + direction.T * distance
    dist = np.sqrt(np.sum(pos**2, axis=1))
    vh = 0.7 * 100 * dist
    vel = xyz.dot(rot) + pos / dist * vh + (direction * vpec).T
    np.testing.assert_allclose(sky["xyz_mpc"], pos, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(sky["vxyz_kms"], vel, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ra", [0.0, 30.0, -30.0])
@pytest.mark.parametrize("dec", [0.0, 30.0, -30.0])
def test_sky_location(ra, dec):
    """tests/test_sources.py:375-395."""
    sky = O.sky_coordinates(np.zeros((1, 3)), np.zeros((1, 3)), 3.0, ra, dec)
    assert sky["ra_deg"][0] == pytest.approx(ra % 360.0, abs=1e-10)
    assert sky["dec_deg"][0] == pytest.approx(dec, abs=1e-10)


def test_tan_projection_properties():
    """Gnomonic identities: the tangent point maps to the reference pixel, a
    meridian through it to a vertical line, and angular offsets to tan(theta)."""
    cube = O.CubeSpec(100, 100, 4, 36.0, 1.0, ra_deg=40.0, dec_deg=25.0)
    p = O.world_to_pixel([40.0], [25.0], [0.0], cube)
    assert p[0, 0] == pytest.approx(49.5, abs=1e-12) and p[1, 0] == pytest.approx(49.5, abs=1e-12)
    p = O.world_to_pixel(np.full(3, 40.0), [24.0, 25.5, 26.0], np.zeros(3), cube)
    np.testing.assert_allclose(p[0], 49.5, atol=1e-9)
    # along the meridian the offset is tan(ddec) in degrees of the tangent plane
    np.testing.assert_allclose(p[1] - 49.5, np.rad2deg(np.tan(np.deg2rad([-1.0, 0.5, 1.0]))) * 100,
                               rtol=1e-12)


def test_channel_edges_bandwidth_and_linearity():
    """tests/test_datacube.py:95-98 and :331-371."""
    for freq in (False, True):
        w = 4.0 if not freq else 18920.0
        cube = O.CubeSpec(16, 16, 16, 15.0, w, spectral_centre=210.0, freq_mode=freq)
        e = cube.channel_edges()
        assert abs(e[-1] - e[0]) == pytest.approx(16 * (w * 1e3 if not freq else w), rel=1e-14)
        assert np.all(np.abs(np.diff(np.diff(e))) < 1e-5)
        ve = cube.velocity_channel_edges()
        assert np.all(np.diff(ve) < 0)  # velocity decreases with channel index
        assert np.all(np.abs(np.diff(np.diff(ve))) < 1e-5)


def test_thermal_sigma():
    """tests/test_spectral_models.py:43-48."""
    assert O.thermal_sigma_kms(1e4) == pytest.approx(9.0853727258, rel=1e-10)


@pytest.mark.parametrize("kind,sigma", [("gauss", "thermal"), ("gauss", 7.0), ("dirac", 0.0)])
def test_spectra_flux(kind, sigma):
    """tests/test_spectral_models.py:15-34 and :96-115: at 1 Mpc the spectrum
    integrates to mHI / 2.36e5 Jy km/s."""
    cube = O.CubeSpec(256, 256, 64, 15.0, 4.0, spectral_centre=70.0)
    sky = O.sky_coordinates(np.full((1, 3), 1e-3), np.zeros((1, 3)), 1.0)
    sig = O.thermal_sigma_kms(1e4) if sigma == "thermal" else sigma
    S = O.spectra(sky["vlos_kms"], np.array([sig]), cube.velocity_channel_edges(), [1e4],
                  sky["distance_mpc"], kind=kind)
    flux = S[0].sum() * 4.0
    assert flux == pytest.approx(1e4 / 2.36e5 / sky["distance_mpc"][0] ** 2, rel=1e-5)


def test_profile_sums_to_one():
    """tests/test_spectral_models.py:50-72."""
    cube = O.CubeSpec(16, 16, 64, 15.0, 4.0, spectral_centre=210.0)
    S = O.spectra([210.0], [9.0853727258], cube.velocity_channel_edges(), [2.36e5], [1.0])
    assert (S[0] * 4.0).sum() == pytest.approx(1.0, rel=1e-4)


def test_spectrum_is_erf_difference():
    """Closed form against scipy.special.erf directly at one channel."""
    cube = O.CubeSpec(4, 4, 8, 10.0, 3.0, spectral_centre=100.0)
    e = cube.velocity_channel_edges() / 1e3
    v, s = 101.3, 2.2
    S = O.spectra([v], [s], cube.velocity_channel_edges(), [2.36e5], [1.0])
    c = 3
    lo, hi = e[c + 1], e[c]
    phi = 0.5 * (erf((hi - v) / (np.sqrt(2) * s)) - erf((lo - v) / (np.sqrt(2) * s)))
    assert S[0, c] * 3.0 == pytest.approx(phi, rel=1e-14)


def test_spectra_nonmonotonic_edges_raise():
    with pytest.raises(ValueError, match="not monotonic"):
        O.spectra([0.0], [1.0], np.array([0.0, 1.0, 0.5]), [1.0], [1.0])


TRUTH = [(0.0, True), (3.0, True), (9.0, False), (-3.0, True), (-9.0, False)]
VTRUTH = [(0.0, True), (3.0, True), (7.0, False), (-3.0, True), (-7.0, False)]


@pytest.mark.parametrize("ra_off,ra_in", TRUTH)
@pytest.mark.parametrize("dec_off,dec_in", TRUTH)
@pytest.mark.parametrize("v_off,v_in", VTRUTH)
@pytest.mark.parametrize("mass_off,mass_in", [(0, False), (1, True)])
def test_prune_truth_table(ra_off, ra_in, dec_off, dec_in, v_off, v_in, mass_off, mass_in):
    """tests/test_martini.py:330-443 with all three prunes on: 1 kpc = 1 arcsec,
    a 2x2x2 cube padded by 5 px, cubic kernel (h = 1 px -> r = 2), Gaussian line
    of 1 km/s (4 channels)."""
    D = KPC_PER_ARCSEC_DIST
    xyz = np.full((1, 3), 1e-3)
    sky = O.sky_coordinates(xyz, np.zeros((1, 3)), D, ra_off / 3600.0, dec_off / 3600.0, h=0.7,
                            vpec_kms=v_off)
    cube = O.CubeSpec(2, 2, 2, 1.0, 1.0, spectral_centre=0.7 * 100 * D, padx=5, pady=5)
    pix = O.world_to_pixel(sky["ra_deg"], sky["dec_deg"], sky["vlos_kms"], cube)
    K = O.KernelSpec("cubic")
    hpx = O.sm_lengths_px([1.0], sky["distance_mpc"], 1.0)
    r = np.ceil(hpx * K.size_in_fwhm)
    keep = O.prune_mask(pix, r, [mass_off * 1e4], cube, 1.0)
    assert bool(keep[0]) == all((ra_in, dec_in, v_in, mass_in))


def test_prune_nan():
    """tests/test_martini.py:445-493: a particle behind the tangent plane gets
    NaN pixel coordinates and is pruned."""
    xyz = np.array([[0.0, 0.0, 0.0], [-20000.0, 0.0, 0.0]])
    sky = O.sky_coordinates(xyz, np.zeros((2, 3)), 10.0)
    cube = O.CubeSpec(2, 2, 2, 10.0, 1.0, spectral_centre=700.0, ra_deg=180.0)
    pix = O.world_to_pixel(sky["ra_deg"], sky["dec_deg"], sky["vlos_kms"], cube)
    assert np.isnan(pix[0, 0]) and np.isnan(pix[1, 0])
    keep = O.prune_mask(pix, np.ones(2), [1e6, 1e6], cube, 0.0, spectral=False)
    assert list(keep) == [False, True]


def single_particle(kernel, spectrum, hsm, sigma=7.0, pad=16):
    """tests/conftest.py:23-122 single particle, dc_zeros-like 16^3 cube with padding."""
    cube = O.CubeSpec(16, 16, 16, 15.0, 4.0, spectral_centre=210.0, padx=pad, pady=pad)
    return O.prepare(np.full((1, 3), 1e-3), np.zeros((1, 3)), [1e4], [hsm], [1e4],
                     distance_mpc=3.0, cube=cube, kernel=kernel, spectrum=spectrum, sigma=sigma)


SIMPLE = [("wc2", None), ("wc6", None), ("cubic", None), ("quartic", None), ("gauss", 3.0),
          ("dirac", None)]


@pytest.mark.parametrize("name,t", SIMPLE)
@pytest.mark.parametrize("spectrum", ["dirac", "gauss"])
def test_mass_accuracy(name, t, spectrum):
    """tests/test_martini.py:205-241: mass recovered within 1% after insertion."""
    K = O.KernelSpec(name, truncate=t)
    P = single_particle(K, spectrum, 0.1 if name == "dirac" else 1.0)
    cube = O.to_jy_per_arcsec2(cube_c.cube(P), P.cube)
    assert O.inserted_mass(cube, P.cube, 3.0) == pytest.approx(1e4, rel=1e-2)


@pytest.mark.parametrize("name,t", SIMPLE)
def test_c_cube_matches_pixel_loop(name, t):
    """The particle-ordered C cube equals the reference's pixel loop."""
    rng = np.random.default_rng(11)
    n = 60
    xyz = (rng.random((n, 3)) - 0.5) * 10
    vxyz = (rng.random((n, 3)) - 0.5) * 40
    hsm = rng.uniform(0.05, 0.4, n) if name == "dirac" else rng.uniform(0.8, 2.0, n)
    cube = O.CubeSpec(16, 16, 16, 15.0, 4.0, spectral_centre=210.0, padx=2, pady=3)
    K = O.KernelSpec(name, truncate=t)
    P = O.prepare(xyz, vxyz, np.full(n, 1e4), hsm, rng.uniform(1e3, 1e4, n), distance_mpc=3.0,
                  cube=cube, kernel=K)
    a = O.cube_pixel_loop(P)
    b = cube_c.cube(P, nthreads=3)
    assert a.max() > 0
    # WendlandC6's closed form cancels ~1e3x near its edge (operation order matters there)
    atol = (1e-10 if name == "wc6" else 1e-13) * a.max()
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=atol)


def test_c_cube_adaptive_matches_pixel_loop():
    rng = np.random.default_rng(12)
    n = 80
    xyz = (rng.random((n, 3)) - 0.5) * 10
    vxyz = (rng.random((n, 3)) - 0.5) * 40
    hsm = np.exp(rng.uniform(np.log(0.05), np.log(3.0), n))
    cube = O.CubeSpec(20, 18, 12, 15.0, 4.0, spectral_centre=210.0)
    for pref in ("cubic", "wc2"):
        P = O.prepare(xyz, vxyz, np.full(n, 1e4), hsm, np.full(n, 1e4), distance_mpc=3.0,
                      cube=cube, kernel=O.adaptive(O.KernelSpec(pref)))
        assert set(np.unique(P.kernel_indices)) == {0, 1, 2}
        a = O.cube_pixel_loop(P)
        b = cube_c.cube(P, nthreads=2)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-13 * a.max())


def test_c_cube_thread_count_invariant():
    rng = np.random.default_rng(13)
    n = 500
    P = O.prepare((rng.random((n, 3)) - 0.5) * 10, (rng.random((n, 3)) - 0.5) * 40,
                  np.full(n, 1e4), np.full(n, 1.0), np.full(n, 1e4), distance_mpc=3.0,
                  cube=O.CubeSpec(16, 16, 16, 15.0, 4.0, spectral_centre=210.0),
                  kernel=O.KernelSpec("cubic"))
    a = cube_c.cube(P, nthreads=1)
    b = cube_c.cube(P, nthreads=4)
    np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-15 * a.max())


def _windmill():
    """The reference's four-armed windmill (tests/test_sources.py:36-45): 200
    particles per arm in the x-y plane, rotating right-handed about z, with
    speeds growing outwards and unit masses; its angular momentum points
    along +z."""
    arm_xyz = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0]], float)
    arm_v = np.array([[0, 1, 0], [-1, 0, 0], [0, -1, 0], [1, 0, 0]], float)
    k = np.arange(1, 201, dtype=float).reshape(200, 1, 1)
    return np.vstack(arm_xyz * k), np.vstack(arm_v * k), np.ones(800)


@pytest.mark.parametrize("axis,target", [("z", [0, 0, 1]), ("x", [1, 0, 0]), ("y", [0, 1, 0])])
def test_L_align_windmill(axis, target):
    """tests/test_sources.py:36-75 of the reference: the rotation takes the
    windmill's angular momentum (+z) to the requested axis.  The reference
    checks both input layouts: (3, N) coordinates give a matrix acting on
    column vectors (rot @ z), (N, 3) coordinates its transpose acting on row
    vectors (z @ rot.T); the oracle returns the (3, N) form.  A wrong axis
    roll sends z to another axis and fails."""
    xyz, vxyz, m = _windmill()
    rot = O.L_align(xyz, vxyz, m, Laxis=axis)
    np.testing.assert_allclose(rot @ np.array([0.0, 0.0, 1.0]), target, atol=1e-12)
    np.testing.assert_allclose(np.array([[0.0, 0.0, 1.0]]) @ rot.T, [target], atol=1e-12)
    np.testing.assert_allclose(rot @ rot.T, np.eye(3), atol=1e-12)
    assert np.linalg.det(rot) == pytest.approx(1.0, abs=1e-12)


def test_L_align_product_matches_windmill():
    """The product's L_align (x axis, as SPHSource.rotate uses it) sends the
    windmill's angular momentum to +x, as the oracle's does."""
    from martini_amd.sources import _L_align

    xyz, vxyz, m = _windmill()
    rot = _L_align(xyz, vxyz, m)
    np.testing.assert_allclose(rot @ np.array([0.0, 0.0, 1.0]), [1.0, 0.0, 0.0], atol=1e-12)
    np.testing.assert_allclose(rot, O.L_align(xyz, vxyz, m, Laxis="x"), atol=1e-13)


@pytest.mark.parametrize("incl", [0.0, 30.0, -30.0])
@pytest.mark.parametrize("az_rot", [0.0, 30.0, -30.0])
@pytest.mark.parametrize("pa", [270.0, 300.0, 240.0])
def test_rotate_L_coords(incl, az_rot, pa):
    """tests/test_sources.py:465-501 of the reference: the L_coords rotation is
    L_align (to x), then az_rot about x, incl about y, and the position angle
    about x, offset by 90 deg for incl >= 0 and 270 deg otherwise (composed here
    with scipy's rotations).  Independently of that composition, the windmill's
    angular momentum must end at (cos|i|, -sin|i| cos pa, -sin|i| sin pa): at
    |incl| from the line of sight (x), projected on the sky along the position
    angle the same way for either sign of incl."""
    from scipy.spatial.transform import Rotation

    from martini_amd import L_coords, SPHSource

    xyz, vxyz, m = _windmill()
    rot = O.L_coords_matrix(xyz, vxyz, m, incl, az_rot, pa)
    want = Rotation.from_euler("x", np.deg2rad(az_rot)).as_matrix() @ O.L_align(
        xyz, vxyz, m, Laxis="x")
    want = Rotation.from_euler("y", np.deg2rad(incl)).as_matrix() @ want
    off = pa - 90.0 if incl >= 0 else pa - 270.0
    want = Rotation.from_euler("x", np.deg2rad(off)).as_matrix() @ want
    np.testing.assert_allclose(rot, want, atol=1e-13)
    i, p = np.deg2rad(abs(incl)), np.deg2rad(pa)
    np.testing.assert_allclose(rot @ np.array([0.0, 0.0, 1.0]),
                               [np.cos(i), -np.sin(i) * np.cos(p), -np.sin(i) * np.sin(p)],
                               atol=1e-12)
    src = SPHSource(distance=3.0, L_coords=L_coords(incl=incl, az_rot=az_rot, pa=pa),
                    T_g=np.full(800, 1e4), mHI_g=m, xyz_g=xyz, vxyz_g=vxyz,
                    hsm_g=np.full(800, 1.0))
    np.testing.assert_allclose(src.current_rotation, rot, atol=1e-13)
