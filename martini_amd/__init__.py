"""martini_amd: MARTINI's source-insertion path, native to AMD MI355X (gfx950).

Drop-in surface of ``martini`` for ``Martini(source, datacube, sph_kernel,
spectral_model, ...).insert_source_in_cube()`` and the SPH-kernel and
spectral-model plugin classes.  FP64 setup runs on the host; binning, the
kernel and line integrals, and the insertion contraction run in hand-written
HIP kernels reached through a C ABI (include/martini_mi355x.h).
"""

__version__ = "0.1.0"

from .datacube import DataCube, HIfreq
from .L_coords import L_coords
from .martini import GlobalProfile, Martini
from .sources import SPHSource
from . import beams, noise, sph_kernels, spectral_models

__all__ = ["Martini", "GlobalProfile", "DataCube", "SPHSource", "L_coords", "HIfreq", "beams",
           "noise", "sph_kernels", "spectral_models"]
