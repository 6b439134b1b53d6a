"""Sag (surface shape) operators.

Each class mirrors one of ``optika.sags``: same name, same fields, same
meaning.  Lengths are millimetres.  The sag, normal and intercept arithmetic
runs inside the fused trace kernel (``csrc/trace_kernel.hip``); the Python
classes carry parameters and expose the reference's per-operator protocol
(``__call__``, ``normal``, ``intercept``, ``propagate_rays``) by tracing
through a one-surface system on the device.

Reference: ``optika/sags/_abc.py:13-122``, ``_flat.py:12-64``,
``_spherical.py:39-191``, ``_conic.py:15-252``, ``_parabolic.py:13-158``,
``_cylindrical.py:13-160``, ``_toroidal.py:13-88``.
"""

from __future__ import annotations

import dataclasses
import math

from .transformations import AbstractTransformation

__all__ = [
    "AbstractSag",
    "NoSag",
    "SphericalSag",
    "AbstractConicSag",
    "ConicSag",
    "ParabolicSag",
    "CylindricalSag",
    "ToroidalSag",
]

# kind codes shared with include/optk.h (OPTK_SAG_*)
SAG_NONE = 0
SAG_SPHERE = 1
SAG_CONIC = 2
SAG_PARABOLA = 3
SAG_CYLINDER = 4
SAG_TOROID = 5


@dataclasses.dataclass(eq=False)
class AbstractSag:
    """Base class of every sag profile (``optika/sags/_abc.py:13-45``)."""

    transformation: None | AbstractTransformation = dataclasses.field(
        default=None, kw_only=True
    )
    """Transformation from the sag frame into the surface frame."""

    parameters_slope_error: object = dataclasses.field(default=None, kw_only=True)
    parameters_roughness: object = dataclasses.field(default=None, kw_only=True)
    parameters_microroughness: object = dataclasses.field(default=None, kw_only=True)

    kind: int = dataclasses.field(default=SAG_NONE, init=False, repr=False)

    def _params(self) -> tuple[float, float, float, float]:
        """Scalar parameters packed into the device record."""
        return (0.0, 0.0, 0.0, 0.0)

    # -- per-operator protocol, evaluated on the device -------------------
    def __call__(self, position):
        from . import _ops

        return _ops.sag_value(self, position)

    def normal(self, position):
        from . import _ops

        return _ops.sag_normal(self, position)

    def intercept(self, rays):
        from . import _ops

        return _ops.sag_intercept(self, rays)

    def propagate_rays(self, rays):
        from . import _ops

        return _ops.sag_propagate(self, rays)


@dataclasses.dataclass(eq=False)
class NoSag(AbstractSag):
    """A flat sag profile, ``z = 0`` (``optika/sags/_flat.py:12-64``)."""

    def __post_init__(self):
        self.kind = SAG_NONE


@dataclasses.dataclass(eq=False)
class SphericalSag(AbstractSag):
    """``z = c r^2 / (1 + sqrt(1 - c^2 r^2))`` (``optika/sags/_spherical.py:56-191``)."""

    radius: float = math.inf

    def __post_init__(self):
        self.kind = SAG_SPHERE

    @property
    def curvature(self) -> float:
        return 1 / self.radius

    def _params(self):
        return (float(self.radius), 0.0, 0.0, 0.0)


@dataclasses.dataclass(eq=False)
class AbstractConicSag(AbstractSag):
    """A conic of revolution (``optika/sags/_conic.py:15-170``)."""

    @property
    def radius(self) -> float:
        raise NotImplementedError

    @property
    def conic(self) -> float:
        raise NotImplementedError

    def _params(self):
        return (float(self.radius), float(self.conic), 0.0, 0.0)


@dataclasses.dataclass(eq=False)
class ConicSag(AbstractConicSag):
    """``z = c r^2 / (1 + sqrt(1 - (1 + k) c^2 r^2))`` (``optika/sags/_conic.py:173-252``)."""

    radius: float = math.inf
    conic: float = 0.0

    def __post_init__(self):
        self.kind = SAG_CONIC


@dataclasses.dataclass(eq=False)
class ParabolicSag(AbstractConicSag):
    """A paraboloid with its own normal and intercept formulas
    (``optika/sags/_parabolic.py:13-158``)."""

    focal_length: float = math.inf

    def __post_init__(self):
        self.kind = SAG_PARABOLA

    @property
    def radius(self) -> float:
        return 2 * self.focal_length

    @property
    def conic(self) -> float:
        return -1

    def _params(self):
        return (float(self.radius), -1.0, float(self.focal_length), 0.0)


@dataclasses.dataclass(eq=False)
class CylindricalSag(AbstractSag):
    """A cylinder with its axis along y (``optika/sags/_cylindrical.py:13-160``)."""

    radius: float = math.inf

    def __post_init__(self):
        self.kind = SAG_CYLINDER

    def _params(self):
        return (float(self.radius), 0.0, 0.0, 0.0)


@dataclasses.dataclass(eq=False)
class ToroidalSag(AbstractSag):
    """``z = R - sqrt((R - z_y)^2 - x^2)`` with ``z_y`` the circle sag of ``y``
    (``optika/sags/_toroidal.py:13-88``).

    The reference finds the intercept with a secant root finder
    (``optika/sags/_abc.py:76-107``); the kernel uses a Newton iteration on
    the same residual, iterated to convergence in float64.
    """

    radius: float = math.inf
    """The minor radius."""

    radius_of_rotation: float = 0.0
    """The major radius."""

    def __post_init__(self):
        self.kind = SAG_TOROID

    def _params(self):
        return (float(self.radius), float(self.radius_of_rotation), 0.0, 0.0)
