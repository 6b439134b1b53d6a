"""Distortion models fitted to traced rays: the map from scene coordinates
(wavelength, field) to sensor coordinates and back.

Mirrors ``optika.distortion.PolynomialDistortionModel``
(``optika/distortion/_distortion.py:276-398``): a polynomial of total degree
``degree`` in (wavelength, field x, field y) for the sensor position
(``distort``) and its fitted inverse in (wavelength, sensor x, sensor y) for
the field (``undistort``), each restricted to the points ``where`` is true.
The fits use ``radiometry.PolynomialFit`` (least squares on coordinates
scaled to [-1, 1]); the reference's ``named_arrays.PolynomialFitFunctionArray``
is not vendored.
"""

from __future__ import annotations

import dataclasses
from typing import Any

from .radiometry import PolynomialFit
from .vectors import Cartesian2dVectorArray

__all__ = ["AbstractDistortionModel", "PolynomialDistortionModel"]


class AbstractDistortionModel:
    """``_distortion.py:25-66``."""

    def distort(self, wavelength, field) -> Cartesian2dVectorArray:
        raise NotImplementedError

    def undistort(self, wavelength, position) -> Cartesian2dVectorArray:
        raise NotImplementedError


@dataclasses.dataclass(eq=False)
class PolynomialDistortionModel(AbstractDistortionModel):
    """Scene coordinates ``wavelength``, ``field_x``, ``field_y`` and the
    sensor coordinates ``sensor_x``, ``sensor_y`` they image to (one shape),
    fitted where ``where`` is true."""

    wavelength: Any = None
    field_x: Any = None
    field_y: Any = None
    sensor_x: Any = None
    sensor_y: Any = None
    degree: int = 2
    where: Any = None

    def __post_init__(self):
        self.fit = PolynomialFit([self.wavelength, self.field_x, self.field_y],
                                 [self.sensor_x, self.sensor_y], self.degree, self.where)
        self.fit_inverse = PolynomialFit([self.wavelength, self.sensor_x, self.sensor_y],
                                         [self.field_x, self.field_y], self.degree, self.where)

    def distort(self, wavelength, field) -> Cartesian2dVectorArray:
        x, y = self.fit(wavelength, field.x, field.y)
        return Cartesian2dVectorArray(x, y)

    def undistort(self, wavelength, position) -> Cartesian2dVectorArray:
        x, y = self.fit_inverse(wavelength, position.x, position.y)
        return Cartesian2dVectorArray(x, y)
