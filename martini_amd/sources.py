"""SPHSource: particle arrays and their FP64 projection onto the sky.

Mirrors martini.sources.SPHSource (martini/sources/sph_source.py:53-724) for
the insertion path: construction (:171-263), rotation / L_coords (:442-526),
translate and boost (:528-569), sky coordinates (:265-326), pixel
coordinates (:328-362) and masking (:364-393).  Units without astropy:
distance Mpc, velocities km/s, positions and smoothing lengths kpc, masses
Msun, temperatures K, angles degrees.  Only the ICRS frame is supported.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
from scipy.spatial.transform import Rotation

from .L_coords import L_coords as _LC


class SkyCoords(NamedTuple):
    """Per-particle observed coordinates (the fields of the reference's SkyCoord
    that the insertion path reads)."""

    ra: np.ndarray  # deg
    dec: np.ndarray  # deg
    distance: np.ndarray  # Mpc
    radial_velocity: np.ndarray  # km/s

    def __getitem__(self, m):
        if isinstance(m, int):
            return tuple.__getitem__(self, m)
        return SkyCoords(*(np.asarray(f)[m] for f in self))


def _L_align(xyz, vxyz, m, frac=0.3):
    """Rotation taking the angular momentum of the innermost ``frac`` of the mass
    to the x axis (martini/sources/_L_align.py:13-120, Laxis='x')."""
    r2 = np.einsum("ij,ij->i", xyz, xyz)
    order = np.argsort(r2, kind="quicksort")
    L = np.cross(xyz, m[:, None] * vxyz)[order]
    cum = np.cumsum(m[order]) / np.sum(m)
    n_in = int(np.argmin(np.abs(cum - frac)))
    n_in = min(max(n_in, 100), len(m))
    Lvec = L[:n_in].sum(axis=0)
    pole = Lvec / np.linalg.norm(Lvec)
    pole = pole / np.linalg.norm(pole)
    ref = np.ones(3) / np.sqrt(3.0)
    if np.array_equal(pole, ref):
        raise RuntimeError("Angular momentum exactly aligned with arbitrarily chosen vector, "
                           "L_align failed.")
    e1 = ref - ref.dot(pole) * pole
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(pole, e1)
    # rows (pole, e1, e2): the pole becomes the x axis
    return np.vstack((pole, e1, e2))


class SPHSource:
    """HI emitting source made of SPH particles."""

    def __init__(self, *, distance: float, vpeculiar: float = 0.0, rotation=None,
                 L_coords=None, ra: float = 0.0, dec: float = 0.0, h: float = 0.7,
                 T_g=None, mHI_g, xyz_g, vxyz_g, hsm_g=None, coordinate_axis=None,
                 coordinate_frame=None) -> None:
        if isinstance(rotation, dict):
            raise ValueError("rotation must be a scipy Rotation or a 3x3 matrix; use "
                             "L_coords=L_coords(incl, az_rot, pa) for angular-momentum alignment")
        if coordinate_frame not in (None, "icrs", "ICRS"):
            raise NotImplementedError("only the ICRS frame is supported")
        xyz = np.asarray(xyz_g, float)
        vxyz = np.asarray(vxyz_g, float)
        if coordinate_axis is None:
            if xyz.shape[0] == 3 and xyz.shape[1] != 3:
                coordinate_axis = 0
            elif xyz.shape[0] != 3 and xyz.shape[1] == 3:
                coordinate_axis = 1
            elif xyz.shape == (3, 3):
                raise RuntimeError("martini.sources.SPHSource: cannot guess coordinate_axis with "
                                   "shape (3, 3), provide explicitly.")
            else:
                raise RuntimeError("martini.sources.SPHSource: incorrect coordinate shape (not "
                                   "(3, N) or (N, 3)).")
        if xyz.shape != vxyz.shape:
            raise ValueError("martini.sources.SPHSource: xyz_g and vxyz_g must have matching "
                             "shapes.")
        if coordinate_axis == 0:
            xyz, vxyz = xyz.T, vxyz.T
        self._xyz = np.array(xyz, float)  # (N, 3) kpc
        self._vxyz = np.array(vxyz, float)  # (N, 3) km/s
        self.npart = self._xyz.shape[0]
        self.coordinate_frame = "icrs"
        self.h = float(h)
        self.T_g = None if T_g is None else np.broadcast_to(np.asarray(T_g, float), (self.npart,)).copy()
        self.mHI_g = np.broadcast_to(np.asarray(mHI_g, float), (self.npart,)).copy()
        self.input_mass = float(self.mHI_g.sum())
        self.hsm_g = None if hsm_g is None else np.broadcast_to(np.asarray(hsm_g, float), (self.npart,)).copy()
        self.ra, self.dec = float(ra), float(dec)
        self.distance = float(distance)
        self.vpeculiar = float(vpeculiar)
        self.vhubble = self.h * 100.0 * self.distance
        self.vsys = self.vhubble + self.vpeculiar
        self._rotation = np.eye(3)
        self.rotate(rotation=rotation, L_coords=L_coords)
        self.skycoords = None
        self.pixcoords = None

    # ---------------------------------------------------- transformations
    @property
    def xyz_g(self) -> np.ndarray:
        return self._xyz

    @property
    def vxyz_g(self) -> np.ndarray:
        return self._vxyz

    @property
    def current_rotation(self) -> np.ndarray:
        return self._rotation

    def rotate(self, rotation=None, *, L_coords=None):
        """Rotate positions and velocities (sph_source.py:442-526)."""
        given = (rotation is not None) + (L_coords is not None)
        if given == 0:
            return Rotation.identity()
        if given > 1:
            raise ValueError("Multiple rotations in a single call not allowed.")
        if rotation is not None:
            mat = rotation.as_matrix() if hasattr(rotation, "as_matrix") else np.asarray(rotation, float)
        else:
            lc = _LC(*L_coords) if not isinstance(L_coords, _LC) else L_coords
            mat = _L_align(self._xyz, self._vxyz, self.mHI_g)
            mat = Rotation.from_euler("x", np.deg2rad(lc.az_rot)).as_matrix() @ mat
            mat = Rotation.from_euler("y", np.deg2rad(lc.incl)).as_matrix() @ mat
            pa_off = lc.pa - 90.0 if lc.incl >= 0 else lc.pa - 270.0
            mat = Rotation.from_euler("x", np.deg2rad(pa_off)).as_matrix() @ mat
        self._xyz = self._xyz @ mat.T
        self._vxyz = self._vxyz @ mat.T
        self._rotation = mat @ self._rotation
        return Rotation.from_matrix(mat)

    def translate(self, translation_vector) -> None:
        """Offset positions by a (3,) vector in kpc (sph_source.py:528-546)."""
        self._xyz = self._xyz + np.asarray(translation_vector, float).reshape(1, 3)

    def boost(self, boost_vector) -> None:
        """Offset velocities by a (3,) vector in km/s (sph_source.py:548-569)."""
        self._vxyz = self._vxyz + np.asarray(boost_vector, float).reshape(1, 3)

    # ---------------------------------------------------- observed frame
    def _observed(self):
        """Positions (Mpc) and velocities (km/s) as seen from the origin: rotate
        by R_z(ra) R_y(-dec), translate by D along the pointing, add the Hubble
        flow h*100*r and the peculiar boost (sph_source.py:276-296)."""
        ra, dec = np.deg2rad(self.ra), np.deg2rad(self.dec)
        R = (Rotation.from_euler("z", ra) * Rotation.from_euler("y", -dec)).as_matrix()
        point = np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)])
        pos = (self._xyz / 1000.0) @ R.T + self.distance * point
        vel = self._vxyz @ R.T + (self.h * 100.0) * pos + self.vpeculiar * point
        return pos, vel

    def _init_skycoords(self, _reset: bool = True) -> None:
        """Sky coordinates (sph_source.py:265-326).  ``_reset`` is accepted for
        signature compatibility; the transformation never alters the stored
        particle arrays."""
        pos, vel = self._observed()
        dist = np.sqrt(np.einsum("ij,ij->i", pos, pos))
        with np.errstate(invalid="ignore", divide="ignore"):
            vlos = np.einsum("ij,ij->i", pos, vel) / dist
        self.skycoords = SkyCoords(
            ra=np.mod(np.rad2deg(np.arctan2(pos[:, 1], pos[:, 0])), 360.0),
            dec=np.rad2deg(np.arctan2(pos[:, 2], np.hypot(pos[:, 0], pos[:, 1]))),
            distance=dist,
            radial_velocity=vlos,
        )

    def _init_pixcoords(self, datacube, origin: int = 0) -> None:
        """Pixel coordinates (3, N), 0-based (sph_source.py:328-362)."""
        if self.skycoords is None:
            raise RuntimeError("Initialize source.skycoords before calling _init_pixcoords.")
        sc = self.skycoords
        self.pixcoords = datacube.world_to_pixel(sc.ra, sc.dec, sc.radial_velocity) + (origin or 0)

    def apply_mask(self, mask) -> None:
        """Keep particles where mask is True (sph_source.py:364-393)."""
        mask = np.asarray(mask)
        if mask.size != self.npart:
            raise ValueError("Mask must have same length as particle arrays.")
        if not mask.any():
            raise RuntimeError("No non-zero mHI source particles in target region.")
        self.npart = int(mask.sum())
        if self.T_g is not None:
            self.T_g = self.T_g[mask]
        self.mHI_g = self.mHI_g[mask]
        self._xyz = self._xyz[mask]
        self._vxyz = self._vxyz[mask]
        if self.hsm_g is not None:
            self.hsm_g = self.hsm_g[mask]
        if self.skycoords is not None:
            self.skycoords = self.skycoords[mask]
        if self.pixcoords is not None:
            self.pixcoords = self.pixcoords[:, mask]
