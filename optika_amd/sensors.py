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
