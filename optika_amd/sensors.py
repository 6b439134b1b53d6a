"""Imaging sensors: the last surface of a system and the binning of rays
into pixels.

Mirrors ``optika.sensors`` (``optika/sensors/_sensors.py:23-500``) for the
trace path: a flat surface whose aperture is the pixel grid
(``width_pixel * num_pixel / 2`` half widths), an ideal material
(``optika/sensors/materials/_materials.py:1566-1616``: vacuum-like, no
refraction) and ``collect``, which histograms rays into
``(wavelength, x, y)`` bins on the device.
"""

from __future__ import annotations

import dataclasses
from typing import Any

from . import apertures as _apertures
from . import materials as _materials
from . import sags as _sags
from .surfaces import AbstractSurface
from .transformations import AbstractTransformation
from .vectors import Cartesian2dVectorArray, as_vector2

__all__ = ["IdealSensorMaterial", "AbstractImagingSensor", "ImagingSensor"]


@dataclasses.dataclass(eq=False)
class IdealSensorMaterial(_materials.Vacuum):
    """Unit quantum efficiency; the refracted cosine is ``-d . n``."""


class AbstractImagingSensor(AbstractSurface):
    @property
    def sag(self):
        return _sags.NoSag()

    @property
    def rulings(self):
        return None

    @property
    def aperture(self):
        """The light-sensitive area (``_sensors.py:83-90``)."""
        w = as_vector2(self.width_pixel)
        n = as_vector2(self.num_pixel)
        return _apertures.RectangularAperture(
            half_width=Cartesian2dVectorArray(w.x * n.x / 2, w.y * n.y / 2)
        )

    def collect(self, rays, wavelength, axis=None, where=True):
        """Bin rays given in sensor coordinates into pixels
        (``_sensors.py:92-171``).

        ``wavelength`` holds the wavelength bin edges (nm).  Returns
        ``(image, direction)``: the intensity-weighted histogram of shape
        ``(n_wavelength, num_pixel.x, num_pixel.y)`` and the flux-weighted mean
        refracted cosine per pixel (1 where a pixel is empty).
        """
        from . import _collect

        return _collect.collect(self, rays, wavelength, where=where)

    def expose(self, image, direction=1, timedelta=None, noise: bool = True, seed: int = 0):
        """Electrons measured from a per-pixel photon-flux image
        (``_sensors.py:173-252``): the flux times the exposure time
        (``timedelta``, default ``timedelta_exposure``), converted by the
        sensor material -- the ideal material's quantum yield is one and its
        noise is Poisson shot noise (``sensors/materials/_materials.py:
        1576-1601``) -- plus zero-mean Gaussian read noise of rms
        ``read_noise``.  Without ``noise`` the expectation.  The random
        numbers are keyed by ``seed`` and the pixel index (the reference's
        ``na.random`` stream is not reproduced).  ``direction`` (the refracted
        cosine) does not change an ideal material's signal."""
        from . import _collect

        if not isinstance(self.material, IdealSensorMaterial):
            raise NotImplementedError("exposure is built for the ideal sensor material only")
        t = self.timedelta_exposure if timedelta is None else timedelta
        return _collect.expose(image, float(t), float(self.read_noise), noise, seed)

    def measure(self, rays, wavelength, where=True, timedelta=None, noise: bool = True,
                seed: int = 0):
        """``collect`` then ``expose`` (``_sensors.py:374-428``): electrons
        per (wavelength bin, pixel)."""
        image, direction = self.collect(rays, wavelength, where=where)
        return self.expose(image, direction, timedelta=timedelta, noise=noise, seed=seed)


@dataclasses.dataclass(eq=False)
class ImagingSensor(AbstractImagingSensor):
    name: None | str = None
    width_pixel: Any = 15e-3
    """Pixel pitch in mm (a scalar or a 2D vector)."""
    axis_pixel: Any = dataclasses.field(
        default_factory=lambda: Cartesian2dVectorArray("detector_x", "detector_y")
    )
    num_pixel: Any = dataclasses.field(default_factory=lambda: Cartesian2dVectorArray(2048, 2048))
    timedelta_exposure: float = 1.0
    material: Any = dataclasses.field(default_factory=IdealSensorMaterial)
    aperture_mechanical: Any = None
    is_field_stop: bool = False
    is_pupil_stop: bool = False
    read_noise: float = 0.0
    transformation: None | AbstractTransformation = None
    kwargs_plot: None | dict = None

    def __post_init__(self):
        self.num_pixel = as_vector2(self.num_pixel)
