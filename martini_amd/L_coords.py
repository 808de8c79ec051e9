"""Source orientation relative to its angular momentum (martini/L_coords.py:7-31).

Angles in degrees.
"""

from typing import NamedTuple


class L_coords(NamedTuple):
    """Inclination, rotation about the angular-momentum axis, and position angle."""

    incl: float = 0.0
    az_rot: float = 0.0
    pa: float = 270.0
