"""Radiometric models fitted to traced rays: the effective area and the
vignetting (relative illumination) of a system.

Mirrors ``optika.radiometry`` (``optika/radiometry/_effective_area.py:13-102``,
``optika/radiometry/_vignetting.py:21-347``).  The reference fits its
polynomials with ``named_arrays.PolynomialFitFunctionArray`` (not vendored);
here the fit is the least-squares polynomial of total degree ``degree`` in the
scene coordinates (wavelength, field x, field y), each coordinate centred and
scaled to [-1, 1] for conditioning, restricted to the points ``where`` is
true.  Host float64 (the fits are small: one value per wavelength and field
point); the rays they are fitted to are traced and reduced on the device.
"""

from __future__ import annotations

import dataclasses
import itertools
from typing import Any

import numpy as np

__all__ = [
    "PolynomialFit",
    "AbstractEffectiveAreaModel",
    "InterpolatedEffectiveAreaModel",
    "AbstractVignettingModel",
    "PolynomialVignettingModel",
]


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


class PolynomialFit:
    """Least-squares polynomial of total degree ``degree`` mapping k input
    coordinates to m outputs."""

    def __init__(self, inputs, outputs, degree: int, where=None):
        x = np.stack([_host(c).reshape(-1) for c in inputs])
        y = np.stack([_host(c).reshape(-1) for c in outputs])
        ok = np.ones(x.shape[1], bool) if where is None else np.asarray(where).reshape(-1)
        ok = ok & np.all(np.isfinite(x), 0) & np.all(np.isfinite(y), 0)
        self.degree = int(degree)
        lo, hi = x[:, ok].min(1), x[:, ok].max(1)
        self.center = (lo + hi) / 2
        half = (hi - lo) / 2
        self.half = np.where(half > 0, half, 1.0)
        self.exponents = [e for e in itertools.product(range(self.degree + 1), repeat=x.shape[0])
                          if sum(e) <= self.degree]
        A = self._design(x[:, ok])
        self.coef, *_ = np.linalg.lstsq(A, y[:, ok].T, rcond=None)

    def _design(self, x):
        u = (x - self.center[:, None]) / self.half[:, None]
        return np.stack([np.prod([u[k] ** e[k] for k in range(len(e))], axis=0)
                         for e in self.exponents], 1)

    def __call__(self, *inputs):
        shape = np.broadcast_shapes(*(np.shape(_host(c)) for c in inputs))
        x = np.stack([np.broadcast_to(_host(c), shape).reshape(-1) for c in inputs])
        y = self._design(x) @ self.coef
        return [y[:, j].reshape(shape) for j in range(y.shape[1])]


class AbstractEffectiveAreaModel:
    """An effective area as a function of wavelength
    (``_effective_area.py:13-34``)."""

    def __call__(self, wavelength):
        raise NotImplementedError


@dataclasses.dataclass(eq=False)
class InterpolatedEffectiveAreaModel(AbstractEffectiveAreaModel):
    """Linear interpolation between calibration points
    (``_effective_area.py:37-102``): ``wavelength`` (nm, ascending) and
    ``area`` (mm^2 for an object at infinity, else deg^2)."""

    wavelength: Any = None
    area: Any = None

    def __call__(self, wavelength):
        return np.interp(_host(wavelength), _host(self.wavelength), _host(self.area))


class AbstractVignettingModel:
    """Relative illumination as a function of the scene coordinates
    (``_vignetting.py:21-58``)."""

    def __call__(self, wavelength, field):
        raise NotImplementedError

    def inverse(self, wavelength, field, radiance):
        """The scene radiance that produced ``radiance`` through the
        vignetting (``_vignetting.py:44-58``)."""
        return _host(radiance) / self(wavelength, field)


@dataclasses.dataclass(eq=False)
class PolynomialVignettingModel(AbstractVignettingModel):
    """A polynomial fit of the illumination over (wavelength, field x,
    field y) (``_vignetting.py:99-185``).  ``wavelength``, ``field_x``,
    ``field_y``, ``illumination`` and ``where`` share one shape."""

    wavelength: Any = None
    field_x: Any = None
    field_y: Any = None
    illumination: Any = None
    degree: int = 2
    where: Any = None

    def __post_init__(self):
        self.fit = PolynomialFit([self.wavelength, self.field_x, self.field_y],
                                 [self.illumination], self.degree, self.where)

    def __call__(self, wavelength, field):
        return self.fit(wavelength, field.x, field.y)[0]
