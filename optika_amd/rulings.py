"""Diffraction-grating rulings: spacing fields, the grating equation and
groove-efficiency models.

Mirrors ``optika.rulings`` (``optika/rulings/_rulings.py:24-1073``,
``optika/rulings/_spacing.py:14-328``).  Spacings and positions are in
millimetres except for :class:`HolographicRulingSpacing`, whose spacing is in
the unit of its recording ``wavelength`` (nanometres), exactly as the
reference's unit algebra makes it.  Groove depths are in nanometres.

The grating equation (:func:`incident_effective`) and every efficiency
model are evaluated per ray inside the fused trace kernel.
"""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np

from .transformations import AbstractTransformation
from .vectors import Cartesian3dVectorArray, as_vector3

__all__ = [
    "incident_effective",
    "AbstractRulingSpacing",
    "ConstantRulingSpacing",
    "Polynomial1dRulingSpacing",
    "HolographicRulingSpacing",
    "AbstractRulings",
    "Rulings",
    "MeasuredRulings",
    "SinusoidalRulings",
    "SquareRulings",
    "SawtoothRulings",
    "TriangularRulings",
    "RectangularRulings",
]

# kind codes shared with include/optk.h (OPTK_SPACING_*, OPTK_RULINGS_*)
SPACING_CONSTANT = 0
SPACING_POLY1D = 1
SPACING_HOLOGRAPHIC = 2

RULINGS_NONE = 0
RULINGS_IDEAL = 1
RULINGS_MEASURED = 2
RULINGS_SINUSOIDAL = 3
RULINGS_SQUARE = 4
RULINGS_SAWTOOTH = 5
RULINGS_TRIANGULAR = 6
RULINGS_RECTANGULAR = 7

MAX_POLY_TERMS = 8


@dataclasses.dataclass(eq=False)
class AbstractRulingSpacing:
    """The local ruling vector ``d * g`` on a grating (``_spacing.py:14-41``)."""

    kind: int = dataclasses.field(default=SPACING_CONSTANT, init=False, repr=False)

    #: factor converting nanometres into this spacing's length unit
    wavelength_scale: float = dataclasses.field(default=1e-6, init=False, repr=False)

    def __call__(self, position, normal):
        from . import _ops

        return _ops.ruling_spacing(self, position, normal)


@dataclasses.dataclass(eq=False)
class ConstantRulingSpacing(AbstractRulingSpacing):
    """``constant * normal`` (``_spacing.py:44-74``)."""

    constant: float = 0.0
    normal: Any = dataclasses.field(default_factory=lambda: Cartesian3dVectorArray(1, 0, 0))

    def __post_init__(self):
        self.kind = SPACING_CONSTANT
        self.wavelength_scale = 1e-6
        self.normal = as_vector3(self.normal)


@dataclasses.dataclass(eq=False)
class Polynomial1dRulingSpacing(AbstractRulingSpacing):
    """``(sum_k c_k x^k) * normal`` with ``x = T(p) . normal`` (``_spacing.py:77-128``).

    ``coefficients`` maps each power to its coefficient (mm per mm^power);
    terms are summed in the dictionary's order, as the reference does.
    """

    coefficients: dict = dataclasses.field(default_factory=dict)
    normal: Any = dataclasses.field(default_factory=lambda: Cartesian3dVectorArray(1, 0, 0))
    transformation: None | AbstractTransformation = None

    def __post_init__(self):
        self.kind = SPACING_POLY1D
        self.wavelength_scale = 1e-6
        self.normal = as_vector3(self.normal)
        if len(self.coefficients) > MAX_POLY_TERMS:
            raise ValueError(
                f"at most {MAX_POLY_TERMS} polynomial terms are supported, "
                f"got {len(self.coefficients)}"
            )


@dataclasses.dataclass(eq=False)
class HolographicRulingSpacing(AbstractRulingSpacing):
    """Rulings recorded by two point sources (``_spacing.py:131-328``)."""

    x1: Any = dataclasses.field(default_factory=Cartesian3dVectorArray)
    x2: Any = dataclasses.field(default_factory=Cartesian3dVectorArray)
    wavelength: float = 500.0
    is_diverging_1: bool = True
    is_diverging_2: bool = False
    transformation: None | AbstractTransformation = None

    def __post_init__(self):
        self.kind = SPACING_HOLOGRAPHIC
        self.wavelength_scale = 1.0
        self.x1 = as_vector3(self.x1)
        self.x2 = as_vector3(self.x2)


def _as_spacing(spacing) -> AbstractRulingSpacing:
    """``AbstractRulings.spacing_`` (``_rulings.py:156-168``)."""
    if isinstance(spacing, AbstractRulingSpacing):
        return spacing
    return ConstantRulingSpacing(constant=float(spacing), normal=(1, 0, 0))


def _orders(order) -> tuple[int, ...]:
    if np.ndim(order) == 0:
        return (int(order),)
    return tuple(int(o) for o in np.asarray(order).ravel())


@dataclasses.dataclass(eq=False)
class AbstractRulings:
    """A ruled surface (``_rulings.py:131-221``).

    ``diffraction_order`` may be an int or a sequence of ints; a sequence adds
    a diffraction-order axis to the traced rays (one output path per order),
    as broadcasting a named order axis does in the reference.
    """

    kind: int = dataclasses.field(default=RULINGS_NONE, init=False, repr=False)

    @property
    def spacing_(self) -> AbstractRulingSpacing:
        return _as_spacing(self.spacing)

    @property
    def orders(self) -> tuple[int, ...]:
        return _orders(self.diffraction_order)

    def _params(self) -> tuple[float, float]:
        return (0.0, 0.0)

    def incident_effective(self, rays, normal):
        from . import _ops

        return _ops.rulings_incident_effective(self, rays, normal)

    def efficiency(self, rays, normal):
        from . import _ops

        return _ops.rulings_efficiency(self, rays, normal)


@dataclasses.dataclass(eq=False)
class Rulings(AbstractRulings):
    """Ideal rulings, efficiency 1 in every order (``_rulings.py:224-251``)."""

    spacing: Any = None
    diffraction_order: Any = 1

    def __post_init__(self):
        self.kind = RULINGS_IDEAL


@dataclasses.dataclass(eq=False)
class MeasuredRulings(AbstractRulings):
    """Efficiency interpolated in wavelength from a measurement
    (``_rulings.py:254-313``).  ``efficiency_measured`` is a pair
    ``(wavelength_nm, efficiency)`` of 1-D arrays."""

    spacing: Any = None
    diffraction_order: Any = 1
    efficiency_measured: Any = None

    def __post_init__(self):
        self.kind = RULINGS_MEASURED


@dataclasses.dataclass(eq=False)
class SinusoidalRulings(AbstractRulings):
    """``J_i(2 gamma)`` (unsquared, as in ``_rulings.py:404-457``)."""

    spacing: Any = None
    depth: float = 0.0
    diffraction_order: Any = 1

    def __post_init__(self):
        self.kind = RULINGS_SINUSOIDAL

    def _params(self):
        return (float(self.depth), 0.0)


@dataclasses.dataclass(eq=False)
class SquareRulings(AbstractRulings):
    """Square-wave grooves (``_rulings.py:548-614``)."""

    spacing: Any = None
    depth: float = 0.0
    diffraction_order: Any = 1

    def __post_init__(self):
        self.kind = RULINGS_SQUARE

    def _params(self):
        return (float(self.depth), 0.0)


@dataclasses.dataclass(eq=False)
class SawtoothRulings(AbstractRulings):
    """Sawtooth grooves (``_rulings.py:705-758``)."""

    spacing: Any = None
    depth: float = 0.0
    diffraction_order: Any = 1

    def __post_init__(self):
        self.kind = RULINGS_SAWTOOTH

    def _params(self):
        return (float(self.depth), 0.0)


@dataclasses.dataclass(eq=False)
class TriangularRulings(AbstractRulings):
    """Triangle-wave grooves (``_rulings.py:849-911``)."""

    spacing: Any = None
    depth: float = 0.0
    diffraction_order: Any = 1

    def __post_init__(self):
        self.kind = RULINGS_TRIANGULAR

    def _params(self):
        return (float(self.depth), 0.0)


@dataclasses.dataclass(eq=False)
class RectangularRulings(AbstractRulings):
    """Rectangular grooves with a duty cycle (``_rulings.py:1008-1073``)."""

    spacing: Any = None
    depth: float = 0.0
    ratio_duty: float = 0.5
    diffraction_order: Any = 1

    def __post_init__(self):
        self.kind = RULINGS_RECTANGULAR

    def _params(self):
        return (float(self.depth), float(self.ratio_duty))


def incident_effective(
    wavelength,
    direction,
    index_refraction,
    normal,
    diffraction_order: int,
    spacing_rulings,
    normal_rulings,
):
    """The effective incident direction on a grating (``_rulings.py:24-128``):
    ``a + sign(a . u) m w g / (n d)``.

    Evaluated on the device for torch inputs; ``wavelength`` and
    ``spacing_rulings`` must be in the same length unit.
    """
    from . import _ops

    return _ops.incident_effective_fn(
        wavelength,
        direction,
        index_refraction,
        normal,
        diffraction_order,
        spacing_rulings,
        normal_rulings,
    )
