"""Rigid coordinate transformations between system and surface frames.

The reference takes these from ``named_arrays.transformations``
(``Cartesian3dTranslation``, ``Cartesian3dRotationX/Y/Z``,
``TransformationList``) and applies them to rays through
``RayVectorArray.__array_matmul__`` / ``__array_add__``
(``optika/rays/_ray_vectors.py:105-166``): a rotation acts on position and
direction, a translation on position only.

Conventions (the reference's third-party library is not importable here, so
these are fixed from the reference's own geometry, see
``optika/systems/_sequential.py:1853-1912``):

* rotations are standard right-handed active rotations, angles in degrees;
  ``RotationY(t) = [[c, 0, s], [0, 1, 0], [-s, 0, c]]``;
* a ``TransformationList`` applies its elements in list order (the first
  element acts first), so ``[RotationY(135), Translation(z=100)]`` rotates
  and then translates;
* the inverse of a list is the reversed list of element inverses.

For the device kernels each transformation is flattened on the host into one
affine map ``p -> R p + t`` (``affine()``); the inverse map is applied as
``R^T (p - t)``.  A map with no rotation is flagged so that the kernel adds
the translation only, which keeps infinite coordinates (rays that miss a
surface) free of ``0 * inf`` products, exactly as a pure translation does.
"""

from __future__ import annotations

import dataclasses
from typing import Sequence

import numpy as np

__all__ = [
    "AbstractTransformation",
    "Cartesian3dTranslation",
    "Cartesian3dRotationX",
    "Cartesian3dRotationY",
    "Cartesian3dRotationZ",
    "TransformationList",
    "Affine",
]


@dataclasses.dataclass(frozen=True)
class Affine:
    """A flattened transformation ``p -> R p + t`` (float64)."""

    matrix: np.ndarray
    vector: np.ndarray
    has_rotation: bool

    @staticmethod
    def identity() -> "Affine":
        return Affine(np.eye(3), np.zeros(3), False)

    def then(self, other: "Affine") -> "Affine":
        """The map that applies ``self`` first and ``other`` second."""
        if not other.has_rotation:
            return Affine(self.matrix, self.vector + other.vector, self.has_rotation)
        if not self.has_rotation:
            return Affine(other.matrix, other.matrix @ self.vector + other.vector, True)
        return Affine(
            other.matrix @ self.matrix,
            other.matrix @ self.vector + other.vector,
            True,
        )


class AbstractTransformation:
    """A rigid transformation from a local frame into its parent frame."""

    def elements(self) -> list["AbstractTransformation"]:
        """The primitive transformations, in application order."""
        return [self]

    @property
    def inverse(self) -> "AbstractTransformation":
        raise NotImplementedError

    def affine(self) -> Affine:
        """Flatten into a single affine map (applied in list order)."""
        result = Affine.identity()
        for e in self.elements():
            result = result.then(e._affine_primitive())
        return result

    def _affine_primitive(self) -> Affine:
        raise NotImplementedError

    def __matmul__(self, other: "AbstractTransformation") -> "TransformationList":
        """``a @ b`` applies ``b`` first, then ``a``."""
        return TransformationList([other, self])


@dataclasses.dataclass
class Cartesian3dTranslation(AbstractTransformation):
    """Translate positions by ``(x, y, z)`` millimetres."""

    x: float = 0.0
    y: float = 0.0
    z: float = 0.0

    @property
    def vector(self) -> np.ndarray:
        return np.array([self.x, self.y, self.z], dtype=np.float64)

    @property
    def inverse(self) -> "Cartesian3dTranslation":
        return Cartesian3dTranslation(-self.x, -self.y, -self.z)

    def _affine_primitive(self) -> Affine:
        return Affine(np.eye(3), self.vector, False)


def _cos_sin_deg(angle: float) -> tuple[float, float]:
    a = np.deg2rad(np.float64(angle))
    return float(np.cos(a)), float(np.sin(a))


@dataclasses.dataclass
class _Rotation(AbstractTransformation):
    angle: float = 0.0
    """Rotation angle in degrees (right-handed, active)."""

    @property
    def inverse(self):
        return type(self)(-self.angle)

    @property
    def matrix(self) -> np.ndarray:
        raise NotImplementedError

    def _affine_primitive(self) -> Affine:
        return Affine(self.matrix, np.zeros(3), True)


class Cartesian3dRotationX(_Rotation):
    @property
    def matrix(self) -> np.ndarray:
        c, s = _cos_sin_deg(self.angle)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=np.float64)


class Cartesian3dRotationY(_Rotation):
    @property
    def matrix(self) -> np.ndarray:
        c, s = _cos_sin_deg(self.angle)
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=np.float64)


class Cartesian3dRotationZ(_Rotation):
    @property
    def matrix(self) -> np.ndarray:
        c, s = _cos_sin_deg(self.angle)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=np.float64)


@dataclasses.dataclass
class TransformationList(AbstractTransformation):
    """A sequence of transformations applied in list order."""

    transformations: Sequence[AbstractTransformation] = ()

    def elements(self) -> list[AbstractTransformation]:
        out = []
        for t in self.transformations:
            out.extend(t.elements())
        return out

    @property
    def inverse(self) -> "TransformationList":
        return TransformationList([t.inverse for t in reversed(self.elements())])
