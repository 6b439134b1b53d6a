"""Apertures and obscurations that clip rays.

Mirrors ``optika.apertures`` (``optika/apertures/_apertures.py:31-1160``).
Radii and half-widths in millimetres clip on the ray *position*; a unitless
aperture (``unit="direction"``) clips on the ray *direction*, as the
reference does when the aperture's bounds are dimensionless
(``_apertures.py:82-102``).  Angles are in degrees.

Comparisons are inclusive (``<=``) for circles, ellipses and rectangles;
``inverted`` flips the mask and ``active=False`` forces it true
(``_apertures.py:309-312``).
"""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np

from .transformations import AbstractTransformation
from .vectors import Cartesian2dVectorArray, Cartesian3dVectorArray, as_vector2

__all__ = [
    "AbstractAperture",
    "CircularAperture",
    "CircularSectorAperture",
    "EllipticalAperture",
    "AbstractPolygonalAperture",
    "PolygonalAperture",
    "RectangularAperture",
    "AbstractRegularPolygonalAperture",
    "RegularPolygonalAperture",
    "OctagonalAperture",
    "IsoscelesTrapezoidalAperture",
]

# kind codes shared with include/optk.h (OPTK_APERTURE_*)
APERTURE_NONE = 0
APERTURE_CIRCLE = 1
APERTURE_SECTOR = 2
APERTURE_ELLIPSE = 3
APERTURE_RECTANGLE = 4
APERTURE_POLYGON = 5

MAX_POLYGON_VERTICES = 64


@dataclasses.dataclass(eq=False)
class AbstractAperture:
    """Base class of every aperture (``_apertures.py:31-102``)."""

    samples_wire: int = dataclasses.field(default=101, kw_only=True)
    active: bool = dataclasses.field(default=True, kw_only=True)
    inverted: bool = dataclasses.field(default=False, kw_only=True)
    transformation: None | AbstractTransformation = dataclasses.field(
        default=None, kw_only=True
    )
    kwargs_plot: None | dict = dataclasses.field(default=None, kw_only=True)
    unit: str = dataclasses.field(default="mm", kw_only=True)
    """``"mm"`` clips on position, ``"direction"`` clips on direction."""

    kind: int = dataclasses.field(default=APERTURE_NONE, init=False, repr=False)

    @property
    def clips_direction(self) -> bool:
        if self.unit not in ("mm", "direction"):
            raise ValueError(f"aperture with unit {self.unit} is not supported")
        return self.unit == "direction"

    def _params(self) -> tuple[float, float, float, float]:
        return (0.0, 0.0, 0.0, 0.0)

    def _vertices_xy(self) -> None | np.ndarray:
        return None

    def __call__(self, position):
        from . import _ops

        return _ops.aperture_mask(self, position)

    def _wire_local(self, num: int) -> np.ndarray:
        raise NotImplementedError(f"{type(self).__name__} has no wire")

    def wire(self, num: None | int = None) -> Cartesian3dVectorArray:
        """``num`` points on the aperture's outline in surface coordinates,
        the aperture's transformation applied (``_apertures.py:121-130``).
        Host float64 arrays."""
        num = self.samples_wire if num is None else int(num)
        p = self._wire_local(num)
        if self.transformation is not None:
            a = self.transformation.affine()
            p = a.matrix @ p + a.vector[:, None]
        return Cartesian3dVectorArray(p[0], p[1], p[2])

    def clip_rays(self, rays):
        from . import _ops

        return _ops.aperture_clip_rays(self, rays)


@dataclasses.dataclass(eq=False)
class CircularAperture(AbstractAperture):
    """``|p.xy| <= radius`` (``_apertures.py:255-314``)."""

    radius: float = 0.0

    def __post_init__(self):
        self.kind = APERTURE_CIRCLE

    def _params(self):
        return (float(self.radius), 0.0, 0.0, 0.0)

    def _wire_local(self, num):
        """Azimuths 0..360 degrees inclusive (``_apertures.py:338-355``)."""
        az = np.deg2rad(np.linspace(0.0, 360.0, num))
        r = float(self.radius)
        return np.stack([r * np.cos(az), r * np.sin(az), np.zeros(num)])

    @property
    def bound_lower(self):
        return Cartesian3dVectorArray(-self.radius, -self.radius, 0.0)

    @property
    def bound_upper(self):
        return Cartesian3dVectorArray(self.radius, self.radius, 0.0)


@dataclasses.dataclass(eq=False)
class CircularSectorAperture(AbstractAperture):
    """A circular sector between two angles (``_apertures.py:395-480``)."""

    radius: float = 0.0
    angle_start: float = 0.0
    angle_stop: float = 180.0

    def __post_init__(self):
        self.kind = APERTURE_SECTOR

    def _params(self):
        return (
            float(self.radius),
            float(np.deg2rad(np.float64(self.angle_start))),
            float(np.deg2rad(np.float64(self.angle_stop))),
            0.0,
        )

    def _wire_local(self, num):
        """Three segments -- the arm out from the vertex, the arc, the arm
        back -- sharing the points evenly, the last one closing the outline
        (``_apertures.py:548-590``)."""
        a0, a1 = np.deg2rad(np.float64(self.angle_start)), np.deg2rad(np.float64(self.angle_stop))
        r0 = float(self.radius)
        per = num / 3
        parts, done = [], 0
        for seg in range(3):
            k = int((seg + 1) * per - done)
            done += k
            t = np.linspace(0.0, 1.0, k, endpoint=done == num)
            if seg == 0:
                r, a = r0 * t, np.full(k, a0)
            elif seg == 1:
                r, a = np.full(k, r0), a0 + (a1 - a0) * t
            else:
                r, a = r0 * (1 - t), np.full(k, a1)
            parts.append(np.stack([r * np.cos(a), r * np.sin(a), np.zeros(k)]))
        return np.concatenate(parts, axis=1)


@dataclasses.dataclass(eq=False)
class EllipticalAperture(AbstractAperture):
    """``(x/rx)^2 + (y/ry)^2 <= 1`` (``_apertures.py:603-663``)."""

    radius: Any = dataclasses.field(default_factory=lambda: Cartesian2dVectorArray(0, 0))

    def __post_init__(self):
        self.kind = APERTURE_ELLIPSE
        self.radius = as_vector2(self.radius)

    def _params(self):
        return (float(self.radius.x), float(self.radius.y), 0.0, 0.0)

    def _wire_local(self, num):
        """Azimuths 0..360 degrees inclusive (``_apertures.py:704-720``)."""
        az = np.deg2rad(np.linspace(0.0, 360.0, num))
        return np.stack([float(self.radius.x) * np.cos(az), float(self.radius.y) * np.sin(az),
                         np.zeros(num)])


@dataclasses.dataclass(eq=False)
class AbstractPolygonalAperture(AbstractAperture):
    """A polygon tested with the crossing-number rule (``_apertures.py:724-778``)."""

    def __post_init__(self):
        self.kind = APERTURE_POLYGON

    @property
    def vertices(self) -> Cartesian3dVectorArray:
        raise NotImplementedError

    def _vertices_xy(self) -> np.ndarray:
        v = self.vertices
        xy = np.stack(
            np.broadcast_arrays(
                np.asarray(v.x, dtype=np.float64), np.asarray(v.y, dtype=np.float64)
            ),
            axis=-1,
        )
        if xy.ndim != 2 or len(xy) > MAX_POLYGON_VERTICES:
            raise ValueError(
                f"polygon must have at most {MAX_POLYGON_VERTICES} vertices"
            )
        return xy

    def _wire_local(self, num):
        """The sides in vertex order, each from its vertex toward the next,
        sharing the points evenly; the last side ends on the first vertex
        (``_apertures.py:794-830``)."""
        v = self.vertices
        xyz = np.stack(np.broadcast_arrays(np.asarray(v.x, np.float64), np.asarray(v.y, np.float64),
                                           np.asarray(v.z, np.float64)))
        nv = xyz.shape[1]
        per = num / nv
        parts, done = [], 0
        for i in range(nv):
            k = int((i + 1) * per - done)
            done += k
            t = np.linspace(0.0, 1.0, k, endpoint=done == num)
            left, right = xyz[:, i], xyz[:, (i + 1) % nv]
            parts.append(left[:, None] + (right - left)[:, None] * t[None, :])
        return np.concatenate(parts, axis=1)

    @property
    def bound_lower(self):
        xy = self._vertices_xy()
        return Cartesian3dVectorArray(xy[:, 0].min(), xy[:, 1].min(), 0.0)

    @property
    def bound_upper(self):
        xy = self._vertices_xy()
        return Cartesian3dVectorArray(xy[:, 0].max(), xy[:, 1].max(), 0.0)


@dataclasses.dataclass(eq=False)
class PolygonalAperture(AbstractPolygonalAperture):
    """A polygon with explicit vertices (``_apertures.py:805-857``)."""

    vertices: Any = None

    def __post_init__(self):
        super().__post_init__()
        v = self.vertices
        if not isinstance(v, Cartesian3dVectorArray):
            v = np.asarray(v, dtype=np.float64)
            z = v[:, 2] if v.shape[1] > 2 else np.zeros(len(v))
            v = Cartesian3dVectorArray(v[:, 0], v[:, 1], z)
        self.vertices = v


@dataclasses.dataclass(eq=False)
class RectangularAperture(AbstractPolygonalAperture):
    """``-hw <= p.xy <= hw`` (``_apertures.py:859-986``)."""

    half_width: Any = 0.0

    def __post_init__(self):
        super().__post_init__()
        self.kind = APERTURE_RECTANGLE
        self.half_width = as_vector2(self.half_width)

    def _params(self):
        return (float(self.half_width.x), float(self.half_width.y), 0.0, 0.0)

    @property
    def vertices(self) -> Cartesian3dVectorArray:
        """The reference builds the corners at 45 + 90 k degrees on a radius
        of sqrt(2), scaled by the half widths (``_apertures.py:970-986``)."""
        az = np.deg2rad(np.linspace(0, 360, 4, endpoint=False) + 45)
        r = np.sqrt(2)
        return Cartesian3dVectorArray(
            r * np.cos(az) * self.half_width.x,
            r * np.sin(az) * self.half_width.y,
            np.zeros(4),
        )

    @property
    def bound_lower(self):
        return Cartesian3dVectorArray(-self.half_width.x, -self.half_width.y, 0.0)

    @property
    def bound_upper(self):
        return Cartesian3dVectorArray(self.half_width.x, self.half_width.y, 0.0)


@dataclasses.dataclass(eq=False)
class AbstractRegularPolygonalAperture(AbstractPolygonalAperture):
    """Vertices at ``radius * (cos a, sin a)``, ``a = 360 k / n`` degrees
    (``_apertures.py:989-1027``)."""

    @property
    def num_vertices(self) -> int:
        raise NotImplementedError

    @property
    def vertices(self) -> Cartesian3dVectorArray:
        angle = np.deg2rad(np.linspace(0, 360, self.num_vertices, endpoint=False))
        return Cartesian3dVectorArray(
            self.radius * np.cos(angle),
            self.radius * np.sin(angle),
            np.zeros(self.num_vertices),
        )


@dataclasses.dataclass(eq=False)
class RegularPolygonalAperture(AbstractRegularPolygonalAperture):
    radius: float = 0.0
    num_vertices: int = 0


@dataclasses.dataclass(eq=False)
class OctagonalAperture(AbstractRegularPolygonalAperture):
    radius: float = 0.0

    @property
    def num_vertices(self) -> int:
        return 8


@dataclasses.dataclass(eq=False)
class IsoscelesTrapezoidalAperture(AbstractPolygonalAperture):
    """A trapezoid symmetric about x (``_apertures.py:1082-1160``)."""

    x_left: float = 0.0
    x_right: float = 0.0
    angle: float = 0.0

    @property
    def vertices(self) -> Cartesian3dVectorArray:
        m = np.tan(np.deg2rad(self.angle) / 2)
        xl, xr = float(self.x_left), float(self.x_right)
        x = np.array([xl, xr, xr, xl])
        y = np.array([m * xl, m * xr, -(m * xr), -(m * xl)])
        return Cartesian3dVectorArray(x, y, np.zeros(4))
