"""FP64 numpy restatement of MARTINI's source-insertion algorithm (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by ``martini_amd``.

Every function cites the reference file:line it restates (paths relative to
the astromartini 2.1.18 source tree).  Units follow the reference's defaults
with astropy removed: kpc for particle positions and smoothing lengths, Mpc for
distances, km/s for velocities, Msun for masses, K for temperatures, arcsec for
pixel sizes, degrees for sky angles, m/s (or Hz) for WCS spectral values.

How each piece is pinned (the checks live in ``tests/test_oracle_*.py``):

* projection (``sky_coordinates``, ``world_to_pixel``): the reference's
  ``test_init_pixcoords`` line-up (tests/test_sources.py:335-373), the
  ``test_ra_dec_rotation`` / ``test_dist_vpec_translation`` matrices
  (:229-316), ``test_sky_location`` (:379-395), and the textbook gnomonic
  identities (a great circle through the tangent point maps to a line).
* channel edges: bandwidth = Nc*width exactly (tests/test_datacube.py:95-98)
  and second differences < 1e-5 m/s (:331-371).
* kernel integrals: line-of-sight quadrature of the 3D kernel (exact for
  WendlandC2, quartic and the cubic's I1/I2; the cubic's I3 carries a known
  +R2*A/8 offset that the normalisation absorbs), unit integral over the plane,
  the R=0 constants, and the reference's own mass-conservation and
  validation-threshold tests (tests/test_sph_kernels.py:44-278).
  WendlandC6 uses the reference's rounded coefficients: pinned to the exact
  line-of-sight integral within 1e-3 relative and by the reference tests only.
* adaptive kernel choice: [0,0,2,1] and [0,2,2,1] (test_sph_kernels.py:284-333).
* spectra: flux sums (tests/test_spectral_models.py:15-34), thermal sigma
  9.0853727258 km/s at 1e4 K (:43-48), unit sum of the profile (:50-72) and
  scipy.special.erf closed form.
* pruning: the reference's truth table (tests/test_martini.py:330-443) and NaN
  case (:445-493).
* cube: mass recovered within 1% (tests/test_martini.py:20-241) and equality
  of the pixel loop with the particle-ordered C cube (``cube_oracle.c``).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf
from scipy.optimize import brentq

# ---------------------------------------------------------------- constants
HI_FREQ_HZ = 1.420405751e9  # martini/datacube.py:21
C_LIGHT_MS = 299792458.0  # astropy.constants.c (exact SI)
K_B = 1.380649e-23  # astropy.constants.k_B (CODATA 2018, exact SI)
M_P = 1.67262192369e-27  # astropy.constants.m_p (CODATA 2018)
ARCSEC_PER_RAD = 180.0 / np.pi * 3600.0
MHI_TO_JY = 2.36e5  # martini/spectral_models.py:119-145 (Msun / (Mpc^2 Jy km/s))


# ---------------------------------------------------------------- rotations
def euler_matrix(axis: str, angle_rad: float) -> np.ndarray:
    """Active right-handed rotation matrix about one axis.

    Same matrices as ``scipy.spatial.transform.Rotation.from_euler(axis, a)
    .as_matrix()``, which the reference uses in sph_source.py:284-285, :489-513.
    """
    c, s = np.cos(angle_rad), np.sin(angle_rad)
    if axis == "x":
        return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])
    if axis == "y":
        return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
    if axis == "z":
        return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    raise ValueError(axis)


def L_align(xyz: np.ndarray, vxyz: np.ndarray, m: np.ndarray, frac: float = 0.3,
            Laxis: str = "x") -> np.ndarray:
    """Rotation matrix putting the angular momentum of the inner ``frac`` of the
    mass along ``Laxis``.  Restates martini/sources/_L_align.py:13-120 as
    SPHSource.rotate calls it (sph_source.py:486-493): coordinates passed as
    (3, N), so the matrix comes back untransposed and acts on column vectors.
    Here ``xyz``/``vxyz`` are (N, 3)."""
    xyz = np.asarray(xyz, float).T
    vxyz = np.asarray(vxyz, float).T
    m = np.asarray(m, float)
    order = np.argsort(np.sum(xyz**2, axis=0), kind="quicksort")
    L = np.cross(xyz, m[np.newaxis] * vxyz, axis=0)[:, order]
    mcum = np.cumsum(m[order]) / np.sum(m)
    nfrac = int(np.argmin(np.abs(mcum - frac)))
    nfrac = min(max(nfrac, 100), len(m))
    Lsum = L[:, :nfrac].sum(axis=1)
    zhat = Lsum / np.sqrt(np.sum(Lsum**2))
    zhat = zhat / np.sqrt(np.sum(zhat**2))
    guess = np.ones(3) / np.sqrt(3.0)
    if (zhat == guess).all():
        raise RuntimeError("Angular momentum exactly aligned with arbitrarily chosen vector")
    xhat = guess - guess.dot(zhat) * zhat
    xhat = xhat / np.sqrt(np.sum(xhat**2))
    yhat = np.cross(zhat, xhat)
    rot = np.vstack((xhat, yhat, zhat))
    shift = {"z": 0, "y": 2, "x": 1}[Laxis]
    return np.roll(rot, shift, axis=0)


def L_coords_matrix(xyz, vxyz, m, incl_deg=0.0, az_rot_deg=0.0, pa_deg=270.0) -> np.ndarray:
    """Composite rotation for ``L_coords`` (martini/sources/sph_source.py:486-513)."""
    rot = L_align(xyz, vxyz, m, frac=0.3, Laxis="x")
    rot = euler_matrix("x", np.deg2rad(az_rot_deg)) @ rot
    rot = euler_matrix("y", np.deg2rad(incl_deg)) @ rot
    pa_off = pa_deg - 90.0 if incl_deg >= 0 else pa_deg - 270.0
    rot = euler_matrix("x", np.deg2rad(pa_off)) @ rot
    return rot


# ----------------------------------------------------------- sky coordinates
def sky_coordinates(xyz_kpc, vxyz_kms, distance_mpc, ra_deg=0.0, dec_deg=0.0,
                    h=0.7, vpec_kms=0.0):
    """Observed position and line-of-sight velocity of each particle.

    Restates SPHSource._init_skycoords (martini/sources/sph_source.py:265-326):
    rotate about y by -dec then about z by ra (:284-285), translate by D*u
    (:286), add per-particle Hubble flow H0*r (:289-296) and the peculiar boost
    vpec*u (:292), then read (ra, dec, distance, radial velocity) of the ICRS
    cartesian representation.  ``xyz_kpc``/``vxyz_kms`` are (N, 3) and already
    carry any source rotation.
    """
    xyz = np.asarray(xyz_kpc, float) / 1000.0  # Mpc
    v = np.asarray(vxyz_kms, float)
    ra, dec = np.deg2rad(ra_deg), np.deg2rad(dec_deg)
    rot = euler_matrix("z", ra) @ euler_matrix("y", -dec)
    unit = np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)])
    r = xyz @ rot.T + distance_mpc * unit
    vel = v @ rot.T + (h * 100.0) * r + vpec_kms * unit
    rho = np.hypot(r[:, 0], r[:, 1])
    dist = np.sqrt(np.sum(r**2, axis=1))
    out = {
        "xyz_mpc": r,
        "vxyz_kms": vel,
        "ra_deg": np.mod(np.rad2deg(np.arctan2(r[:, 1], r[:, 0])), 360.0),
        "dec_deg": np.rad2deg(np.arctan2(r[:, 2], rho)),
        "distance_mpc": dist,
    }
    with np.errstate(invalid="ignore", divide="ignore"):
        out["vlos_kms"] = np.sum(vel * r, axis=1) / dist
    return out


# ---------------------------------------------------------------- data cube
class CubeSpec:
    """Geometry and spectral WCS of a DataCube.

    Restates DataCube.__init__ (martini/datacube.py:157-215) and DataCube.wcs
    (:426-486): crpix = n/2 + 0.5 + pad (1-based), cdelt = (-px, +px, -w) in
    velocity mode or (-px, +px, +df) in frequency mode, crval = (ra, dec,
    spectral centre).  Velocity mode is VRAD in m/s; frequency mode FREQ in Hz;
    the spectral centre is converted with the radio convention
    (datacube.py:202-206).
    """

    def __init__(self, n_px_x, n_px_y, n_channels, px_size_arcsec, channel_width,
                 spectral_centre=0.0, ra_deg=0.0, dec_deg=0.0, padx=0, pady=0,
                 freq_mode=False, centre_is_freq=False):
        self.n_px_x, self.n_px_y, self.n_channels = int(n_px_x), int(n_px_y), int(n_channels)
        self.px_size_arcsec = float(px_size_arcsec)
        self.freq_mode = bool(freq_mode)
        self.channel_width = abs(float(channel_width))  # km/s or Hz
        self.padx, self.pady = int(padx), int(pady)
        self.ra_deg, self.dec_deg = float(ra_deg), float(dec_deg)
        if self.freq_mode:
            # spectral centre as frequency (Hz); velocity input converts radio-wise
            self.crval3 = (float(spectral_centre) if centre_is_freq
                           else HI_FREQ_HZ * (1.0 - float(spectral_centre) * 1e3 / C_LIGHT_MS))
            self.cdelt3 = self.channel_width
        else:
            self.crval3 = (C_LIGHT_MS * (1.0 - float(spectral_centre) / HI_FREQ_HZ)
                           if centre_is_freq else float(spectral_centre) * 1e3)  # m/s
            self.cdelt3 = -self.channel_width * 1e3
        self.crpix = np.array([self.n_px_x / 2.0 + 0.5 + self.padx,
                               self.n_px_y / 2.0 + 0.5 + self.pady,
                               self.n_channels / 2.0 + 0.5])
        self.px_deg = self.px_size_arcsec / 3600.0

    @property
    def shape(self):
        return (self.n_px_x + 2 * self.padx, self.n_px_y + 2 * self.pady, self.n_channels)

    def spectral_world(self, pix0):
        """0-based spectral pixel -> world value (m/s or Hz); linear WCS axis."""
        return self.crval3 + self.cdelt3 * (np.asarray(pix0, float) + 1.0 - self.crpix[2])

    def channel_edges(self):
        """DataCube.channel_edges (martini/datacube.py:513-536): pixels -0.5..Nc-0.5."""
        return self.spectral_world(np.arange(self.n_channels + 1) - 0.5)

    def velocity_channel_edges(self):
        """DataCube.velocity_channel_edges (datacube.py:551-562), m/s, radio."""
        e = self.channel_edges()
        if self.freq_mode:
            return C_LIGHT_MS * (1.0 - e / HI_FREQ_HZ)
        return e


def world_to_pixel(ra_deg, dec_deg, vlos_kms, cube: CubeSpec):
    """0-based pixel coordinates (3, N): TAN projection plus the linear spectral
    axis (SPHSource._init_pixcoords, sph_source.py:328-362, with the WCS of
    datacube.py:426-486; ICRS frame and icrs specsys).  NaN for particles on the
    far hemisphere, as wcslib returns."""
    a0, d0 = np.deg2rad(cube.ra_deg), np.deg2rad(cube.dec_deg)
    a, d = np.deg2rad(np.asarray(ra_deg, float)), np.deg2rad(np.asarray(dec_deg, float))
    da = a - a0
    den = np.sin(d) * np.sin(d0) + np.cos(d) * np.cos(d0) * np.cos(da)
    with np.errstate(invalid="ignore", divide="ignore"):
        xi = np.rad2deg(np.cos(d) * np.sin(da) / den)
        eta = np.rad2deg((np.sin(d) * np.cos(d0) - np.cos(d) * np.sin(d0) * np.cos(da)) / den)
    bad = den <= 0
    xi = np.where(bad, np.nan, xi)
    eta = np.where(bad, np.nan, eta)
    x = cube.crpix[0] - 1.0 - xi / cube.px_deg
    y = cube.crpix[1] - 1.0 + eta / cube.px_deg
    vlos = np.asarray(vlos_kms, float)
    if cube.freq_mode:
        spec = HI_FREQ_HZ * (1.0 - vlos * 1e3 / C_LIGHT_MS)
    else:
        spec = vlos * 1e3
    z = cube.crpix[2] - 1.0 + (spec - cube.crval3) / cube.cdelt3
    return np.vstack((x, y, z))


# ------------------------------------------------------------------ kernels
def W_cubic(q):
    """Cubic spline, sph_kernels.py:758-789 (support 1, FWHM-free units)."""
    q = np.asarray(q, float)
    w = np.where(q < 0.5, 1 - 6 * q**2 + 6 * q**3, 2 * (1 - q) ** 3)
    w = np.where(q > 1, 0.0, w)
    return w * 8 / np.pi


def W_wc2(q):
    """WendlandC2, sph_kernels.py:373-397."""
    q = np.asarray(q, float)
    return np.where(q < 1, (1 - q) ** 4 * (4 * q + 1), 0.0) * 21 / 2 / np.pi


def W_wc6(q):
    """WendlandC6, sph_kernels.py:511-542."""
    q = np.asarray(q, float)
    return np.where(q < 1, (1 - q) ** 8 * (1 + 8 * q + 25 * q**2 + 32 * q**3), 0.0) * 1365 / 64 / np.pi


def W_quartic(q):
    """Quartic spline, sph_kernels.py:1438-1484."""
    q = np.asarray(q, float)
    w = np.where(q < 1, (1 - q) ** 4, 0.0)
    w = w - np.where(q < 0.6, 5 * (0.6 - q) ** 4, 0.0)
    w = w + np.where(q < 0.2, 10 * (0.2 - q) ** 4, 0.0)
    return w * 15625 / 512 / np.pi


GAUSS_SIG = 1.0 / (2.0 * np.sqrt(2.0 * np.log(2.0)))  # sph_kernels.py:1020, FWHM = 1


def W_gauss(q, truncate):
    """Truncated Gaussian, sph_kernels.py:962-992 (with its norm)."""
    q = np.asarray(q, float)
    w = np.where(q < truncate * GAUSS_SIG,
                 (GAUSS_SIG * np.sqrt(2 * np.pi)) ** -3 * np.exp(-(q / GAUSS_SIG) ** 2 / 2), 0.0)
    return w / gauss_norm(truncate)


def gauss_norm(truncate):
    """_GaussianKernel.norm, sph_kernels.py:949-951."""
    t = float(truncate)
    return erf(t / np.sqrt(2)) - 2 * t / np.sqrt(2 * np.pi) * np.exp(-t * t / 2)


def find_fwhm(W) -> float:
    """FWHM of a kernel with W(0) maximal: 2 * root of W(q) - W(0)/2
    (sph_kernels.py:14-48 uses fsolve; brentq on the bracket gives the same root)."""
    half = float(W(np.zeros(1))[0]) / 2
    return 2.0 * brentq(lambda q: float(W(np.array([q]))[0]) - half, 1e-6, 1.0 - 1e-9,
                        xtol=1e-17, rtol=8.9e-16, maxiter=200)


class KernelSpec:
    """One concrete kernel's constants: the reference's kernel classes reduced
    to data (sph_kernels.py:340-1600)."""

    def __init__(self, name, truncate=None, dirac_size=1.0):
        self.name = name
        self.truncate = truncate
        self.min_valid = None
        self.max_valid = None
        if name == "cubic":
            fwhm = find_fwhm(W_cubic)
            self.rescale = self.size_in_fwhm = 1.0 / fwhm
            self.min_valid = 1.16  # :749
        elif name == "wc2":
            fwhm = find_fwhm(W_wc2)
            self.rescale = self.size_in_fwhm = 1.0 / fwhm
            self.min_valid = 1.51  # :362
        elif name == "wc6":
            fwhm = find_fwhm(W_wc6)
            self.rescale = self.size_in_fwhm = 1.0 / fwhm
            self.min_valid = 1.29  # :502
        elif name == "quartic":
            fwhm = find_fwhm(W_quartic)
            self.rescale = self.size_in_fwhm = 1.0 / fwhm
            self.min_valid = 1.2385  # :1429
        elif name == "gauss":
            t = float(truncate)
            if t < 2:
                raise RuntimeError("GaussianKernel with truncation <2sigma will cause large errors in total mass.")
            # min sizes by truncation band, sph_kernels.py:938-947
            self.min_valid = (3.7 if t < 3 else 2.3357 if t < 4 else 1.1288 if t < 5
                              else 0.45 if t < 6 else 0.336)
            self.rescale = 1.0
            self.size_in_fwhm = t * GAUSS_SIG  # :953
            self.norm = gauss_norm(t)
        elif name == "dirac":
            self.rescale = 1.0
            self.size_in_fwhm = float(dirac_size)
            self.max_valid = 0.5  # :1104
        else:
            raise ValueError(name)

    def valid(self, h):
        """``_validate``'s boolean (e.g. sph_kernels.py:460, :1193)."""
        h = np.asarray(h, float)
        if self.max_valid is not None:
            return h <= self.max_valid
        return h >= self.min_valid

    def integral(self, dx, dy, h):
        """``_kernel_integral`` of this kernel: weight per unit pixel area at
        offsets (dx, dy) = particle - pixel centre, effective size ``h`` (px)."""
        return KERNEL_INTEGRALS[self.name](np.asarray(dx, float), np.asarray(dy, float),
                                           np.asarray(h, float), self)


def _ki_cubic(dx, dy, h, spec):
    """_CubicSplineKernel._kernel_integral, sph_kernels.py:791-858."""
    R2 = 4.0 * (dx * dx + dy * dy) / (h * h)
    out = np.zeros(np.broadcast(dx, dy, h).shape)
    out[R2 == 0] = 11.0 / 16.0 + 0.25 * 0.25
    inner = (R2 > 0) & (R2 <= 1)
    outer = (R2 > 1) & (R2 <= 4)
    r = R2[inner]
    A = np.sqrt(1 - r)
    B = np.sqrt(4 - r)
    I1 = (A - 0.5 * A**3 - 1.5 * r * A + 3.0 / 32.0 * A * (3 * r + 2)
          + 9.0 / 32.0 * r * r * (np.log(1 + A) - np.log(np.sqrt(r))))
    I3 = (-B * (3 * r + 56) / 4.0 + A * (4 * r + 50) / 8.0
          - 3.0 / 8.0 * r * (r + 16) * np.log((2 + B) / (1 + A))
          + 2 * (3 * r + 4) * (B - A) + 2 * (B**3 - A**3))
    out[inner] = I1 + 0.25 * I3
    r = R2[outer]
    B = np.sqrt(4 - r)
    I2 = (-B * (3 * r + 56) / 4.0 - 3.0 / 8.0 * r * (r + 16) * np.log((2 + B) / np.sqrt(r))
          + 2 * (3 * r + 4) * B + 2 * B**3)
    out[outer] = 0.25 * I2
    return out / 1.59689476201133 / (h * h) * 4


def _ki_wc2(dx, dy, h, spec):
    """_WendlandC2Kernel._kernel_integral, sph_kernels.py:400-441."""
    R2 = (dx * dx + dy * dy) / (h * h)
    out = np.zeros(np.broadcast(dx, dy, h).shape)
    out[R2 == 0] = 2.0 / 3.0
    use = (R2 < 1) & (R2 != 0)
    r = R2[use]
    A = np.sqrt(1 - r)
    out[use] = (5 * r * r * (0.5 * r + 3) * np.log((1 + A) / np.sqrt(r))
                + A * (-27.0 / 2.0 * r * r - 14.0 / 3.0 * r + 2.0 / 3.0))
    return out * (21 / 2 / np.pi) / (h * h)


# WendlandC6 closed form: the reference's indefinite integral (sph_kernels.py:
# 576-678) as a table of (coefficient, power of R, power of z, factor) with
# factor 0 = plain, 1 = times sqrt(R^2+z^2), 2 = times log(sqrt(R^2+z^2)+z).
# The coefficients are the reference's own (several are rounded), kept so that
# the restatement evaluates the same function.
WC6_TERMS = (
    (-231, 10, 1, 0), (-385, 8, 3, 0), (-1155, 8, 1, 0), (-462, 6, 5, 0),
    (-1540, 6, 3, 0), (-462, 6, 1, 0), (-330, 4, 7, 0), (-1386, 4, 5, 0),
    (-462, 4, 3, 0), (66, 4, 1, 0), (-(128 + 1 / 3), 2, 9, 0), (-660, 2, 7, 0),
    (-277.2, 2, 5, 0), (44, 2, 3, 0), (8 / 3, 0, 11, 1), ((16 + 4 / 15), 2, 9, 1),
    (70.4, 0, 9, 1), (360.8, 2, 7, 1), (132, 0, 7, 1), (550, 2, 5, 1),
    (-11, 2, 1, 0), (7.21875, 12, 0, 2), (24.7813, 10, 1, 1), (173.25, 10, 0, 2),
    (530.75, 8, 1, 1), (288.75, 8, 0, 2), (47.4792, 8, 3, 1), (767.25, 6, 1, 1),
    (58.0167, 6, 5, 1), (819.5, 6, 3, 1), (41.7, 4, 7, 1), (752.4, 4, 5, 1),
    (896.5, 4, 3, 1), (-21, 0, 11, 0), (-(128 + 1 / 3), 0, 9, 0), (-66, 0, 7, 0),
    (13.2, 0, 5, 0), (-11 / 3, 0, 3, 0), (1, 0, 1, 0),
)


def _wc6_indef(R, z):
    S = np.sqrt(R * R + z * z)
    L = np.log(S + z)
    tot = np.zeros(np.broadcast(R, z).shape)
    for c, pr, pz, f in WC6_TERMS:
        t = c * R**pr * z**pz
        if f == 1:
            t = t * S
        elif f == 2:
            t = t * L
        tot = tot + t
    return tot


def _ki_wc6(dx, dy, h, spec):
    """_WendlandC6Kernel._kernel_integral, sph_kernels.py:544-685."""
    R = np.sqrt(dx * dx + dy * dy) / h
    out = np.zeros(np.broadcast(dx, dy, h).shape)
    norm = 1365 / 64 / np.pi
    use = (R < 1) & (R != 0)
    r = R[use]
    zmax = np.sqrt(1 - r * r)
    out[use] = norm * 2 * (_wc6_indef(r, zmax) - _wc6_indef(r, np.zeros_like(r)))
    out[R == 0] = norm * 2 * (4 / 15)
    return out / (h * h)


def _quartic_IA(R, z, A):
    """Indefinite z-integral of (A - q)^4 at fixed R, sph_kernels.py:1518-1552."""
    q = np.sqrt(z * z + R * R)
    R2 = R * R
    return (A**4 * z - 2 * A**3 * z * q + 2 * A**2 * z * (3 * R2 + z * z)
            - A * R2 * (4 * A * A + 3 * R2) * np.arcsinh(z / R) / 2
            - A * z * q * (5 * R2 + 2 * z * z) / 2 + R2 * R2 * z + 2 * R2 * z**3 / 3 + z**5 / 5)


def _ki_quartic(dx, dy, h, spec):
    """_QuarticSplineKernel._kernel_integral, sph_kernels.py:1486-1564."""
    R = np.sqrt(dx * dx + dy * dy) / h
    out = np.zeros(np.broadcast(dx, dy, h).shape)
    for lim, coef in ((0.2, 10.0), (0.6, -5.0), (1.0, 1.0)):
        m = (R > 0) & (R < lim)
        r = R[m]
        out[m] += coef * _quartic_IA(r, np.sqrt(lim * lim - r * r), lim)
    out[R == 0] = 384 / 3125
    out *= 2 * 15625 / 512 / np.pi
    return out / (h * h)


def _ki_gauss(dx, dy, h, spec):
    """_GaussianKernel._kernel_integral, sph_kernels.py:994-1044."""
    t = spec.truncate
    sig = GAUSS_SIG
    dr = np.sqrt(dx * dx + dy * dy)
    with np.errstate(invalid="ignore"):
        zmax = np.sqrt(t * t - (dr / h / sig) ** 2)
    zmax = np.where(t > dr / h / sig, zmax, 0.0)
    c = h * np.sqrt(2) * sig
    x0, x1 = (dx - 0.5) / c, (dx + 0.5) / c
    y0, y1 = (dy - 0.5) / c, (dy + 0.5) / c
    out = 0.25 * erf(zmax / np.sqrt(2)) * (erf(x1) - erf(x0)) * (erf(y1) - erf(y0))
    out = np.where((dr - np.sqrt(0.5)) / h / sig > t, 0.0, out)
    return out / spec.norm


def _ki_dirac(dx, dy, h, spec):
    """DiracDeltaKernel._kernel_integral, sph_kernels.py:1136-1165 (strict <)."""
    return np.where((np.abs(dx) < 0.5) & (np.abs(dy) < 0.5), 1.0, 0.0) * np.ones_like(h)


KERNEL_INTEGRALS = {"cubic": _ki_cubic, "wc2": _ki_wc2, "wc6": _ki_wc6,
                    "quartic": _ki_quartic, "gauss": _ki_gauss, "dirac": _ki_dirac}


class AdaptiveSpec:
    """_AdaptiveKernel (sph_kernels.py:1207-1399): a priority list of kernels;
    the first whose ``valid`` accepts ``h_px * rescale_k`` wins (:1241-1274);
    index -1 falls back to kernels[0]'s size/rescale/integral (:1266-1274,
    :1358-1362)."""

    def __init__(self, kernels):
        self.kernels = tuple(kernels)

    def select(self, h_px):
        h_px = np.asarray(h_px, float)
        idx = -np.ones(h_px.shape, dtype=int)
        for ik, K in enumerate(self.kernels):
            idx[(idx == -1) & K.valid(h_px * K.rescale)] = ik
        return idx

    def per_particle(self, h_px):
        idx = self.select(h_px)
        use = np.where(idx < 0, 0, idx)
        size = np.array([K.size_in_fwhm for K in self.kernels])[use]
        resc = np.array([K.rescale for K in self.kernels])[use]
        return idx, size, resc

    def integral(self, dx, dy, h_eff, idx):
        out = np.zeros(np.broadcast(dx, h_eff).shape)
        use = np.where(idx < 0, 0, idx)
        for ik in np.unique(use):
            m = use == ik
            out[m] = self.kernels[ik].integral(np.asarray(dx)[m], np.asarray(dy)[m],
                                               np.asarray(h_eff)[m])
        return out


def adaptive(preferred: KernelSpec) -> AdaptiveSpec:
    """The public adaptive classes: (preferred, Dirac, Gauss t=6)
    (sph_kernels.py:1631, :1691, :1754, :1825-1831, :1894)."""
    return AdaptiveSpec((preferred, KernelSpec("dirac"), KernelSpec("gauss", truncate=6.0)))


def sm_lengths_px(hsm_kpc, distance_mpc, px_size_arcsec):
    """_init_sm_lengths, sph_kernels.py:235-255: arctan(hsm / d) in pixels."""
    ang = np.arctan(np.asarray(hsm_kpc, float) / (np.asarray(distance_mpc, float) * 1000.0))
    return ang * ARCSEC_PER_RAD / px_size_arcsec


# ------------------------------------------------------------------ spectra
def thermal_sigma_kms(T_K):
    """GaussianSpectrum.half_width, 'thermal' (spectral_models.py:465-485)."""
    return np.sqrt(K_B * np.asarray(T_K, float) / M_P) / 1000.0


def spectra(vlos_kms, sigma_kms, edges_ms, mHI_msun, distance_mpc, kind="gauss"):
    """_BaseSpectrum.init_spectra + evaluate_spectra + spectral_function
    (spectral_models.py:63-200, :347-425, :510-570): N x Nc in Jy.

    Gaussian: 0.5*(erf((hi-v)/sqrt2/sigma) - erf((lo-v)/sqrt2/sigma)).
    Dirac: H(v-lo)*H(hi-v) with H(0) = 1.  Then * mHI/d^2 / |dv| / 2.36e5.
    """
    e = np.asarray(edges_ms, float) / 1000.0  # km/s
    d = np.diff(e)
    if np.all(d > 0):
        lo, hi = e[:-1], e[1:]
    elif np.all(d < 0):
        lo, hi = e[1:], e[:-1]
    else:
        raise ValueError("Channel edges are not monotonic sequence.")
    v = np.asarray(vlos_kms, float)[:, None]
    if kind == "gauss":
        s = np.broadcast_to(np.asarray(sigma_kms, float), np.shape(v)[:1])[:, None]
        phi = 0.5 * (erf((hi[None] - v) / np.sqrt(2.0) / s) - erf((lo[None] - v) / np.sqrt(2.0) / s))
    elif kind == "dirac":
        phi = np.heaviside(v - lo[None], 1.0) * np.heaviside(hi[None] - v, 1.0)
    else:
        raise ValueError(kind)
    A = np.asarray(mHI_msun, float) * np.asarray(distance_mpc, float) ** -2.0
    return A[:, None] * phi / np.abs(d)[None] / MHI_TO_JY


# ------------------------------------------------------------------ pruning
def prune_mask(pixcoords, sm_ranges, mHI, cube: CubeSpec, half_width_kms,
               spatial=True, spectral=True, mass=True):
    """_BaseMartini._prune_particles accept mask (martini/martini.py:168-241)."""
    x, y, z = pixcoords
    r = np.asarray(sm_ranges, float)
    dv_max = np.max(np.abs(np.diff(cube.velocity_channel_edges()))) / 1000.0  # km/s
    hw = np.asarray(half_width_kms, float) / dv_max
    reject = np.zeros(x.shape, bool)
    if spatial:
        with np.errstate(invalid="ignore"):
            reject |= (x + r < 0) | (y + r < 0)
            reject |= x - r > cube.n_px_x + cube.padx * 2
            reject |= y - r > cube.n_px_y + cube.pady * 2
        reject |= np.isnan(x) | np.isnan(y)
    if spectral:
        with np.errstate(invalid="ignore"):
            reject |= z + 4 * hw < 0
            reject |= z - 4 * hw > cube.n_channels
    if mass:
        reject |= np.asarray(mHI) == 0
    return ~reject


# ------------------------------------------------------------ full pipeline
class Prepared:
    """Particle arrays after setup and pruning, ready for insertion."""


def prepare(xyz_kpc, vxyz_kms, mHI, hsm_kpc, T_K, *, distance_mpc, cube: CubeSpec,
            kernel, sigma="thermal", spectrum="gauss", ra_deg=0.0, dec_deg=0.0,
            h=0.7, vpec_kms=0.0, rotation=None, prune=True):
    """Martini.__init__ sequence (martini/martini.py:118-151): sky coords,
    pixel coords, smoothing lengths, ranges, pruning.  ``kernel`` is a
    KernelSpec (non-adaptive) or AdaptiveSpec.  ``sigma`` is 'thermal' or a
    constant in km/s.  ``rotation`` is an optional 3x3 matrix applied first."""
    xyz = np.asarray(xyz_kpc, float)
    vxyz = np.asarray(vxyz_kms, float)
    if rotation is not None:
        xyz = xyz @ np.asarray(rotation).T
        vxyz = vxyz @ np.asarray(rotation).T
    sky = sky_coordinates(xyz, vxyz, distance_mpc, ra_deg, dec_deg, h, vpec_kms)
    pix = world_to_pixel(sky["ra_deg"], sky["dec_deg"], sky["vlos_kms"], cube)
    hpx = sm_lengths_px(hsm_kpc, sky["distance_mpc"], cube.px_size_arcsec)
    n = xyz.shape[0]
    if isinstance(kernel, AdaptiveSpec):
        kidx, size, resc = kernel.per_particle(hpx)
    else:
        kidx = np.zeros(n, int)
        size = np.full(n, kernel.size_in_fwhm)
        resc = np.full(n, kernel.rescale)
    ranges = np.ceil(hpx * size)
    if spectrum == "gauss":
        sig = thermal_sigma_kms(T_K) if isinstance(sigma, str) else np.full(n, float(sigma))
        hw = sig
    else:
        sig = np.zeros(n)
        hw = 0.0
    mHI = np.broadcast_to(np.asarray(mHI, float), (n,)).copy()
    keep = (prune_mask(pix, ranges, mHI, cube, hw) if prune else np.ones(n, bool))
    if not keep.any():
        raise RuntimeError("No non-zero mHI source particles in target region.")
    P = Prepared()
    P.pixcoords = pix[:, keep]
    P.h_px = hpx[keep]
    P.kernel_indices = kidx[keep]
    P.rescale = resc[keep]
    P.size_in_fwhm = size[keep]
    P.sm_ranges = ranges[keep]
    P.vlos_kms = sky["vlos_kms"][keep]
    P.distance_mpc = sky["distance_mpc"][keep]
    P.mHI = mHI[keep]
    P.sigma_kms = sig[keep]
    P.kernel = kernel
    P.spectrum = spectrum
    P.cube = cube
    P.keep = keep
    return P


def particle_spectra(P: Prepared):
    return spectra(P.vlos_kms, P.sigma_kms, P.cube.velocity_channel_edges(), P.mHI,
                   P.distance_mpc, kind=P.spectrum)


def kernel_weights(P: Prepared, dx, dy, sel=Ellipsis):
    """_px_weight (sph_kernels.py:85-119): h_eff = h_px * rescale, then the
    concrete (or adaptive) integral."""
    h_eff = P.h_px[sel] * P.rescale[sel]
    if isinstance(P.kernel, AdaptiveSpec):
        return P.kernel.integral(dx, dy, h_eff, P.kernel_indices[sel])
    return P.kernel.integral(dx, dy, h_eff)


def cube_pixel_loop(P: Prepared, S=None):
    """The reference's pixel loop (martini/martini.py:243-283, :327-343):
    per pixel, square mask over all particles, weights, weighted spectra sum.
    Returns Jy/pix^2 (before the arcsec^2 conversion).  O(Npix * N): small
    cases only."""
    if S is None:
        S = particle_spectra(P)
    nx, ny, nc = P.cube.shape
    out = np.zeros((nx, ny, nc))
    x, y = P.pixcoords[0], P.pixcoords[1]
    for i in range(nx):
        for j in range(ny):
            m = (np.abs(i - x) <= P.sm_ranges) & (np.abs(j - y) <= P.sm_ranges)
            if not m.any():
                continue
            w = kernel_weights(P, x[m] - i, y[m] - j, m)
            out[i, j] = (S[m] * w[:, None]).sum(axis=0)
    return out


def to_jy_per_arcsec2(cube_jy_pix2, cube: CubeSpec):
    """Unit conversion after insertion (martini/martini.py:364-366)."""
    return cube_jy_pix2 / cube.px_size_arcsec**2


def inserted_mass(cube_jy_arcsec2, cube: CubeSpec, distance_mpc):
    """Mass recovered from the cube (tests/test_martini.py:20-56)."""
    F = (cube_jy_arcsec2 * cube.px_size_arcsec**2).sum((0, 1))
    dv = np.abs(np.diff(cube.velocity_channel_edges())) / 1000.0
    return np.sum(MHI_TO_JY * distance_mpc**2 * F * dv)


# ------------------------------------------------------------ beam convolution
def convolve_beam(cube, kernel, padx, pady):
    """Martini.convolve_beam (martini/martini.py:863-901): every channel map of
    the padded cube convolved with the beam image by
    scipy.signal.fftconvolve(..., mode="same"), then the padding dropped
    (DataCube.drop_pad, datacube.py:710-729).  The caller multiplies by the
    beam area to go from Jy/arcsec^2 to Jy/beam.  Pinned by the convolution
    theorem (scipy's FFT) and, in tests, by a direct sum on small maps."""
    from scipy.signal import fftconvolve

    out = np.empty_like(cube, dtype=float)
    for c in range(cube.shape[2]):
        out[:, :, c] = fftconvolve(cube[:, :, c], kernel, mode="same")
    if padx > 0 and pady > 0:
        out = out[padx:-padx, pady:-pady, :]
    return out


# -------------------------------------------------------------- global profile
def global_profile(xyz_kpc, vxyz_kms, mHI, hsm_kpc, T_K, *, distance_mpc, n_channels,
                   channel_width, spectral_centre, sigma="thermal", spectrum="gauss",
                   rotation=None, h=0.7):
    """GlobalProfile.insert_source_in_spectrum (martini/martini.py:1369-1596):
    a 1-pixel cube of 1 degree pixels (datacube.py:889-940), every particle in
    the spectral band contributing its whole line (no spatial aperture; the
    prune keeps spectral and mass cuts only, martini.py:1536-1548), summed over
    particles, in Jy per channel.  Pinned by the flux-sum identity of the
    spectra (sum over channels = mHI / (2.36e5 d^2 |dv|) for lines inside the
    band)."""
    cube = CubeSpec(1, 1, n_channels, 3600.0, channel_width, spectral_centre=spectral_centre)
    P = prepare(xyz_kpc, vxyz_kms, mHI, hsm_kpc, T_K, distance_mpc=distance_mpc, cube=cube,
                kernel=KernelSpec("dirac", dirac_size=np.inf), sigma=sigma, spectrum=spectrum,
                rotation=rotation, h=h, prune=False)
    hw = P.sigma_kms if spectrum == "gauss" else 0.0
    keep = prune_mask(P.pixcoords, P.sm_ranges, P.mHI, cube, hw, spatial=False)
    return particle_spectra(P)[keep].sum(axis=0), int(keep.sum())
