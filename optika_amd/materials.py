"""Optical materials: vacuum, mirrors, glass and multilayer coatings.

Mirrors ``optika.materials`` (``optika/materials/_materials.py:20-455``,
``_layers.py:23-645``, ``_multilayers.py:33-1110``, ``matrices.py:20-246``,
``profiles.py:20-541``, ``_snells_law.py:13-366``).

Units: wavelengths and layer thicknesses in nanometres, attenuation in 1/mm,
Sellmeier ``c`` coefficients in square micrometres, interface widths in
nanometres.

Every material's ``index_refraction``, ``attenuation`` and ``efficiency``
(including the multilayer transfer-matrix reflectivity) and the vector
Snell's law are evaluated per ray inside the fused trace kernel.  The free
functions :func:`snells_law`, :func:`snells_law_scalar` and
:func:`multilayer_efficiency` dispatch to the same device code.
"""

from __future__ import annotations

import dataclasses
from typing import Any, Sequence

from .chemicals import AbstractChemical, Chemical

__all__ = [
    "snells_law_scalar",
    "snells_law",
    "multilayer_coefficients",
    "multilayer_efficiency",
    "AbstractMaterial",
    "Vacuum",
    "AbstractMirror",
    "Mirror",
    "MeasuredMirror",
    "Glass",
    "AbstractLayer",
    "Layer",
    "LayerSequence",
    "PeriodicLayerSequence",
    "AbstractMultilayerMaterial",
    "MultilayerMirror",
    "MultilayerFilm",
    "profiles",
]

# kind codes shared with include/optk.h (OPTK_MATERIAL_*, OPTK_PROFILE_*)
MATERIAL_VACUUM = 0
MATERIAL_MIRROR = 1
MATERIAL_MEASURED_MIRROR = 2
MATERIAL_GLASS = 3
MATERIAL_MULTILAYER_MIRROR = 4
MATERIAL_MULTILAYER_FILM = 5
MATERIAL_SILICON_SENSOR = 6

PROFILE_NONE = 0
PROFILE_ERF = 1
PROFILE_EXPONENTIAL = 2
PROFILE_LINEAR = 3
PROFILE_SINUSOIDAL = 4


class profiles:
    """Interface roughness profiles (``optika/materials/profiles.py``).

    Each scales the Fresnel reflection coefficient of an interface by the
    Fourier transform of the profile's derivative evaluated at
    ``s = Re(4 pi n cos(theta) / wavelength)``.
    """

    @dataclasses.dataclass(eq=False)
    class AbstractInterfaceProfile:
        width: float = 0.0
        kind: int = dataclasses.field(default=PROFILE_NONE, init=False, repr=False)

    @dataclasses.dataclass(eq=False)
    class ErfInterfaceProfile(AbstractInterfaceProfile):
        """``exp(-(s w)^2 / 2)`` (``profiles.py:129-222``)."""

        def __post_init__(self):
            self.kind = PROFILE_ERF

    @dataclasses.dataclass(eq=False)
    class ExponentialInterfaceProfile(AbstractInterfaceProfile):
        """``1 / (1 + (s w)^2 / 2)`` (``profiles.py:225-324``)."""

        def __post_init__(self):
            self.kind = PROFILE_EXPONENTIAL

    @dataclasses.dataclass(eq=False)
    class LinearInterfaceProfile(AbstractInterfaceProfile):
        """``sin(x) / x``, ``x = sqrt(3) w s`` (``profiles.py:327-430``)."""

        def __post_init__(self):
            self.kind = PROFILE_LINEAR

    @dataclasses.dataclass(eq=False)
    class SinusoidalInterfaceProfile(AbstractInterfaceProfile):
        """``pi (sin(x - pi/2)/(x - pi/2) + sin(x + pi/2)/(x + pi/2)) / 4``,
        ``x = a w s``, ``a = pi / (pi^2 - 8)`` (``profiles.py:433-541``)."""

        def __post_init__(self):
            self.kind = PROFILE_SINUSOIDAL


# -- materials ---------------------------------------------------------------


@dataclasses.dataclass(eq=False)
class AbstractMaterial:
    """Interface of every material (``_materials.py:20-78``)."""

    kind: int = dataclasses.field(default=MATERIAL_VACUUM, init=False, repr=False)

    @property
    def transformation(self):
        return None

    @property
    def is_mirror(self) -> bool:
        return False

    def index_refraction(self, rays):
        from . import _ops

        return _ops.material_index_refraction(self, rays)

    def attenuation(self, rays):
        from . import _ops

        return _ops.material_attenuation(self, rays)

    def efficiency(self, rays, normal):
        from . import _ops

        return _ops.material_efficiency(self, rays, normal)


@dataclasses.dataclass(eq=False)
class Vacuum(AbstractMaterial):
    """n = 1, no attenuation, efficiency 1 (``_materials.py:81-116``)."""

    def __post_init__(self):
        self.kind = MATERIAL_VACUUM


@dataclasses.dataclass(eq=False)
class AbstractMirror(AbstractMaterial):
    """Reflects; passes the ray's n and attenuation through (``_materials.py:119-157``)."""

    @property
    def is_mirror(self) -> bool:
        return True


@dataclasses.dataclass(eq=False)
class Mirror(AbstractMirror):
    """A perfect reflector (``_materials.py:160-173``)."""

    substrate: Any = None

    def __post_init__(self):
        self.kind = MATERIAL_MIRROR


@dataclasses.dataclass(eq=False)
class MeasuredMirror(AbstractMirror):
    """Reflectivity interpolated in wavelength (``_materials.py:176-305``).

    ``efficiency_measured`` is a pair ``(wavelength_nm, efficiency)`` of 1-D
    arrays measured at a single incidence angle.
    """

    efficiency_measured: Any = None
    substrate: Any = None
    serial_number: Any = None

    def __post_init__(self):
        self.kind = MATERIAL_MEASURED_MIRROR


@dataclasses.dataclass(eq=False)
class Glass(AbstractMaterial):
    """Three-term Sellmeier glass (``_materials.py:308-455``)."""

    b1: float = 0.0
    b2: float = 0.0
    b3: float = 0.0
    c1: float = 0.0
    c2: float = 0.0
    c3: float = 0.0

    def __post_init__(self):
        self.kind = MATERIAL_GLASS

    @classmethod
    def n_bk7(cls) -> "Glass":
        """SCHOTT N-BK7 (``_materials.py:383-396``)."""
        return cls(
            b1=1.03961212,
            b2=0.231792344,
            b3=1.01046945,
            c1=0.00600069867,
            c2=0.0200179144,
            c3=103.560653,
        )

    @classmethod
    def f2(cls) -> "Glass":
        """SCHOTT F2 (``_materials.py:398-411``)."""
        return cls(
            b1=1.34533359,
            b2=0.209073176,
            b3=0.937357162,
            c1=0.00997743871,
            c2=0.0470450767,
            c3=111.886764,
        )


# -- layers ------------------------------------------------------------------


class AbstractLayer:
    """A slab (or stack of slabs) of optical material (``_layers.py:23-107``)."""


@dataclasses.dataclass(eq=False)
class Layer(AbstractLayer):
    """A homogeneous layer (``_layers.py:134-277``); ``chemical=None`` is vacuum."""

    chemical: None | str | AbstractChemical = None
    thickness: None | float = None
    interface: None | profiles.AbstractInterfaceProfile = None
    kwargs_plot: None | dict = None
    x_label: Any = None

    @property
    def _chemical(self) -> None | AbstractChemical:
        c = self.chemical
        if c is None or isinstance(c, AbstractChemical):
            return c
        return Chemical(c)

    def n(self, wavelength):
        c = self._chemical
        return 1 if c is None else c.n(wavelength)


@dataclasses.dataclass(eq=False)
class LayerSequence(AbstractLayer):
    """An explicit sequence of layers (``_layers.py:381-499``)."""

    layers: Sequence[AbstractLayer] = ()


@dataclasses.dataclass(eq=False)
class PeriodicLayerSequence(AbstractLayer):
    """A period of layers repeated ``num_periods`` times (``_layers.py:527-645``).

    The transfer matrix is built as in the reference: the period once with
    the ambient medium, once more with the closing layer's medium, then that
    second matrix raised to the power ``num_periods - 1`` by binary
    powering.
    """

    layers: Sequence[AbstractLayer] = ()
    num_periods: int = 1


@dataclasses.dataclass(eq=False)
class AbstractMultilayerMaterial(AbstractMaterial):
    """Passes n and attenuation through (``_multilayers.py:771-826``)."""

    @property
    def _layers(self) -> AbstractLayer:
        layers = self.layers
        if isinstance(layers, AbstractLayer):
            return layers
        return LayerSequence(list(layers) if layers is not None else [])


@dataclasses.dataclass(eq=False)
class MultilayerMirror(AbstractMultilayerMaterial):
    """Efficiency = s/p-averaged multilayer reflectivity (``_multilayers.py:901-1110``)."""

    layers: Any = None
    substrate: None | Layer = None

    def __post_init__(self):
        self.kind = MATERIAL_MULTILAYER_MIRROR

    @property
    def is_mirror(self) -> bool:
        return True


@dataclasses.dataclass(eq=False)
class MultilayerFilm(AbstractMultilayerMaterial):
    """Efficiency = s/p-averaged multilayer transmissivity (``_multilayers.py:829-898``)."""

    layers: Any = None

    def __post_init__(self):
        self.kind = MATERIAL_MULTILAYER_FILM

    @property
    def substrate(self) -> None:
        return None


# -- free functions ------------------------------------------------------------


def snells_law_scalar(cos_incidence, index_refraction, index_refraction_new):
    """Cosine of the refracted angle (``_snells_law.py:13-38``), on the device."""
    from . import _ops

    return _ops.snells_law_scalar(cos_incidence, index_refraction, index_refraction_new)


def snells_law(
    direction,
    index_refraction,
    index_refraction_new,
    normal=None,
    is_mirror: bool = False,
):
    """Vector form of Snell's law (``_snells_law.py:41-366``), on the device."""
    from . import _ops

    return _ops.snells_law(direction, index_refraction, index_refraction_new, normal, is_mirror)


def multilayer_coefficients(wavelength, direction=1, n=1, layers=None, substrate=None):
    """Complex ``(r, t)`` of a stack (``_multilayers.py:33-237``), on the device."""
    from . import _ops

    return _ops.multilayer_coefficients(wavelength, direction, n, layers, substrate)


def multilayer_efficiency(wavelength, direction=1, n=1, layers=None, substrate=None):
    """Reflectivity and transmissivity of a stack (``_multilayers.py:240-532``),
    on the device."""
    from . import _ops

    return _ops.multilayer_efficiency(wavelength, direction, n, layers, substrate)

