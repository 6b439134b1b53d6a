"""Small Cartesian vector containers used by the host-side API.

The reference builds every quantity on named-axis arrays
(``named_arrays.Cartesian3dVectorArray`` and friends).  Here a vector is a
thin record of three (or two) components, each of which may be a Python
float, a numpy array or a torch tensor.  Only the operations that the trace
path's public API needs are provided: component access, dot and cross
products, length and normalisation, and elementwise arithmetic.

Units are fixed rather than carried: lengths in millimetres, wavelengths in
nanometres, angles in degrees.

Reference counterparts:
``optika/vectors/_vectors_object.py:136-160`` (ObjectVectorArray),
``optika/vectors/_vectors_polarization.py:36-38`` (PolarizationVectorArray.average).
"""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np

__all__ = [
    "Cartesian2dVectorArray",
    "Cartesian3dVectorArray",
    "ObjectVectorArray",
    "PolarizationVectorArray",
    "as_vector3",
    "as_vector2",
]


def _sqrt(a):
    if hasattr(a, "sqrt") and not isinstance(a, np.ndarray):
        return a.sqrt()
    return np.sqrt(a)


@dataclasses.dataclass
class Cartesian2dVectorArray:
    """A 2D vector with components ``x`` and ``y``."""

    x: Any = 0.0
    y: Any = 0.0

    def __iter__(self):
        return iter((self.x, self.y))

    def __add__(self, other):
        o = as_vector2(other)
        return Cartesian2dVectorArray(self.x + o.x, self.y + o.y)

    def __sub__(self, other):
        o = as_vector2(other)
        return Cartesian2dVectorArray(self.x - o.x, self.y - o.y)

    def __mul__(self, other):
        if isinstance(other, Cartesian2dVectorArray):
            return Cartesian2dVectorArray(self.x * other.x, self.y * other.y)
        return Cartesian2dVectorArray(self.x * other, self.y * other)

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, Cartesian2dVectorArray):
            return Cartesian2dVectorArray(self.x / other.x, self.y / other.y)
        return Cartesian2dVectorArray(self.x / other, self.y / other)

    def __neg__(self):
        return Cartesian2dVectorArray(-self.x, -self.y)

    @property
    def length(self):
        return _sqrt(self.x * self.x + self.y * self.y)


@dataclasses.dataclass
class Cartesian3dVectorArray:
    """A 3D vector with components ``x``, ``y`` and ``z``."""

    x: Any = 0.0
    y: Any = 0.0
    z: Any = 0.0

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    @property
    def xy(self) -> Cartesian2dVectorArray:
        return Cartesian2dVectorArray(self.x, self.y)

    def replace(self, **kwargs) -> "Cartesian3dVectorArray":
        return dataclasses.replace(self, **kwargs)

    def __add__(self, other):
        o = as_vector3(other)
        return Cartesian3dVectorArray(self.x + o.x, self.y + o.y, self.z + o.z)

    __radd__ = __add__

    def __sub__(self, other):
        o = as_vector3(other)
        return Cartesian3dVectorArray(self.x - o.x, self.y - o.y, self.z - o.z)

    def __rsub__(self, other):
        o = as_vector3(other)
        return Cartesian3dVectorArray(o.x - self.x, o.y - self.y, o.z - self.z)

    def __mul__(self, other):
        if isinstance(other, Cartesian3dVectorArray):
            return Cartesian3dVectorArray(
                self.x * other.x, self.y * other.y, self.z * other.z
            )
        return Cartesian3dVectorArray(self.x * other, self.y * other, self.z * other)

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, Cartesian3dVectorArray):
            return Cartesian3dVectorArray(
                self.x / other.x, self.y / other.y, self.z / other.z
            )
        return Cartesian3dVectorArray(self.x / other, self.y / other, self.z / other)

    def __neg__(self):
        return Cartesian3dVectorArray(-self.x, -self.y, -self.z)

    def __matmul__(self, other):
        """Dot product, summed in x, y, z order."""
        o = as_vector3(other)
        return self.x * o.x + self.y * o.y + self.z * o.z

    def cross(self, other) -> "Cartesian3dVectorArray":
        o = as_vector3(other)
        return Cartesian3dVectorArray(
            self.y * o.z - self.z * o.y,
            self.z * o.x - self.x * o.z,
            self.x * o.y - self.y * o.x,
        )

    @property
    def length(self):
        return _sqrt(self.x * self.x + self.y * self.y + self.z * self.z)

    @property
    def normalized(self) -> "Cartesian3dVectorArray":
        return self / self.length

    def to_numpy(self) -> np.ndarray:
        """Stack the components into an array of shape ``(3, ...)``."""
        comps = []
        for c in (self.x, self.y, self.z):
            if hasattr(c, "detach"):
                c = c.detach().cpu().numpy()
            comps.append(np.asarray(c, dtype=np.float64))
        comps = np.broadcast_arrays(*comps)
        return np.stack(comps)


def as_vector3(v) -> Cartesian3dVectorArray:
    """Coerce a vector, a 3-sequence or a scalar into a 3D vector."""
    if isinstance(v, Cartesian3dVectorArray):
        return v
    if isinstance(v, (tuple, list)) and len(v) == 3:
        return Cartesian3dVectorArray(*v)
    if isinstance(v, np.ndarray) and v.shape[:1] == (3,):
        return Cartesian3dVectorArray(v[0], v[1], v[2])
    return Cartesian3dVectorArray(v, v, v)


def as_vector2(v) -> Cartesian2dVectorArray:
    """Coerce a vector, a 2-sequence or a scalar into a 2D vector."""
    if isinstance(v, Cartesian2dVectorArray):
        return v
    if isinstance(v, (tuple, list)) and len(v) == 2:
        return Cartesian2dVectorArray(*v)
    if isinstance(v, np.ndarray) and v.shape[:1] == (2,):
        return Cartesian2dVectorArray(v[0], v[1])
    return Cartesian2dVectorArray(v, v)


@dataclasses.dataclass
class ObjectVectorArray:
    """Input grid of a sequential system: wavelength, field and pupil.

    Mirrors ``optika.vectors.ObjectVectorArray``
    (``optika/vectors/_vectors_object.py:136-160``).  ``field`` is in degrees
    when the object is at infinity and in millimetres otherwise; ``pupil`` the
    other way round.  ``wavelength`` is in nanometres.
    """

    wavelength: Any = 500.0
    field: Any = dataclasses.field(default_factory=Cartesian2dVectorArray)
    pupil: Any = dataclasses.field(default_factory=Cartesian2dVectorArray)

    def __post_init__(self):
        self.field = as_vector2(self.field)
        self.pupil = as_vector2(self.pupil)

    def copy_shallow(self) -> "ObjectVectorArray":
        return dataclasses.replace(self)


@dataclasses.dataclass
class PolarizationVectorArray:
    """A pair of ``s`` and ``p`` polarization components.

    ``average`` follows ``optika/vectors/_vectors_polarization.py:36-38``.
    """

    s: Any = 0.0
    p: Any = 0.0

    @property
    def average(self):
        return (self.s + self.p) / 2
