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

import numpy as np

from . import apertures as _apertures
from . import materials as _materials
from . import sags as _sags
from .surfaces import AbstractSurface
from .transformations import AbstractTransformation
from .vectors import Cartesian2dVectorArray, as_vector2

__all__ = ["IdealSensorMaterial", "BackIlluminatedSiliconSensorMaterial", "AbstractImagingSensor",
           "ImagingSensor", "absorption_effective", "transmittance", "absorbance",
           "charge_collection_efficiency", "quantum_efficiency_effective"]


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


# -- silicon sensor absorbance (optika/sensors/materials/_materials.py:55-376) --

THICKNESS_OXIDE = 5.0        # nm (50 A, Stern 1994; _stern_1994.py:3)
THICKNESS_SUBSTRATE = 7000.0  # nm (7 um, _stern_1994.py:4)
THICKNESS_IMPLANT = 231.7     # nm (2317 A, _stern_1994.py:5)
CCE_BACKSURFACE = 0.21        # _stern_1994.py:7


def _np(v):
    """A host numpy copy of an array or device tensor."""
    return v.detach().cpu().numpy() if hasattr(v, "detach") else np.asarray(v)


def absorption_effective(wavelength, n_substrate, direction_substrate):
    """Absorption coefficient per unit perpendicular depth (1/nm) of a wave
    refracted into the light-sensitive region, ``4 pi Im(n cos) / lambda``
    (``_materials.py:55-93``)."""
    return 4 * np.pi * np.imag(_np(n_substrate) * _np(direction_substrate)) / \
        _np(wavelength).astype(np.float64)


def transmittance(wavelength, direction=1, n=1, thickness_oxide=THICKNESS_OXIDE,
                  thickness_substrate=THICKNESS_SUBSTRATE, chemical_oxide="SiO2",
                  chemical_substrate="Si", roughness_oxide=0.0, roughness_substrate=0.0):
    """Fraction of the incident energy transmitted through the oxide into
    the substrate, per polarization (``_materials.py:96-197``): the
    multilayer transfer-matrix kernel on the device."""
    from . import _ops
    from .materials import Layer, profiles

    oxide = Layer(chemical=chemical_oxide, thickness=float(thickness_oxide),
                  interface=profiles.ErfInterfaceProfile(width=float(roughness_oxide)))
    substrate = Layer(chemical=chemical_substrate, thickness=float(thickness_substrate),
                      interface=profiles.ErfInterfaceProfile(width=float(roughness_substrate)))
    _, t = _ops.multilayer_efficiency(wavelength, direction, n, layers=[oxide],
                                      substrate=substrate)
    return t


def absorbance(wavelength, direction=1, n=1, thickness_oxide=THICKNESS_OXIDE,
               thickness_substrate=THICKNESS_SUBSTRATE, chemical_oxide="SiO2",
               chemical_substrate="Si", roughness_oxide=0.0, roughness_substrate=0.0,
               method: str = "Beer-Lambert"):
    """Fraction of the incident energy absorbed in the light-sensitive
    region, per polarization (``_materials.py:200-376``), by the
    Beer-Lambert method: the transmittance into the substrate times
    ``1 - exp(-alpha d)`` with the effective absorption of the refracted
    wave.  The ``exact`` method (interference inside the substrate) is not
    built."""
    from . import _ops
    from .chemicals import Chemical
    from .vectors import PolarizationVectorArray

    if method != "Beer-Lambert":
        raise NotImplementedError(f"absorbance method {method!r}: only 'Beer-Lambert' is built")
    t = transmittance(wavelength, direction, n, thickness_oxide, thickness_substrate,
                      chemical_oxide, chemical_substrate, roughness_oxide, roughness_substrate)
    chem = chemical_substrate if not isinstance(chemical_substrate, str) else Chemical(chemical_substrate)
    n_sub = chem.n(_np(wavelength).astype(np.float64))
    d_sub = _ops.snells_law_scalar(direction, n, n_sub)
    alpha = absorption_effective(wavelength, n_sub, d_sub)
    decay = np.exp(-alpha * float(thickness_substrate))

    def absorbed(tp):
        tp = _np(tp)
        return np.real(tp - tp * decay)

    return PolarizationVectorArray(absorbed(t.s), absorbed(t.p))


def charge_collection_efficiency(absorption, thickness_implant=THICKNESS_IMPLANT,
                                 cce_backsurface=CCE_BACKSURFACE):
    """Average charge collection efficiency of the piecewise-linear
    differential profile (``_materials.py:377-497``):
    ``eta_0 + (1 - eta_0) (1 - exp(-alpha W)) / (alpha W)`` with the
    absorption per unit perpendicular depth ``alpha`` (1/nm) and the implant
    thickness ``W`` (nm)."""
    z0 = _np(absorption) * thickness_implant
    exp_z0 = np.exp(-z0)
    term_1 = cce_backsurface
    term_2 = ((1 - cce_backsurface) / z0) * (1 - exp_z0)
    return term_1 + term_2


def quantum_efficiency_effective(wavelength, direction=1, n=1, thickness_oxide=THICKNESS_OXIDE,
                                 thickness_implant=THICKNESS_IMPLANT,
                                 thickness_substrate=THICKNESS_SUBSTRATE,
                                 cce_backsurface=CCE_BACKSURFACE, chemical_oxide="SiO2",
                                 chemical_substrate="Si", roughness_oxide=0.0,
                                 roughness_substrate=0.0):
    """Effective quantum efficiency ``absorbance.average * CCE``
    (``_materials.py:500-687``)."""
    from . import _ops
    from .chemicals import Chemical

    a = absorbance(wavelength, direction, n, thickness_oxide, thickness_substrate,
                   chemical_oxide, chemical_substrate, roughness_oxide, roughness_substrate)
    chem = chemical_substrate if not isinstance(chemical_substrate, str) else Chemical(chemical_substrate)
    n_sub = chem.n(_np(wavelength).astype(np.float64))
    d_sub = _ops.snells_law_scalar(direction, n, n_sub)
    cce = charge_collection_efficiency(absorption_effective(wavelength, n_sub, d_sub),
                                       thickness_implant, cce_backsurface)
    return _np(a.average) * cce


def _cos_incidence(direction, normal):
    """``-direction @ normal``, with the reference's defaults (normal
    incidence on a surface facing -z)."""
    from .vectors import Cartesian3dVectorArray, as_vector3

    if direction is None:
        direction = Cartesian3dVectorArray(0.0, 0.0, 1.0)
        normal = Cartesian3dVectorArray(0.0, 0.0, -1.0)
    d, nrm = as_vector3(direction), as_vector3(normal)
    return -(_np(d.x) * _np(nrm.x) + _np(d.y) * _np(nrm.y) + _np(d.z) * _np(nrm.z))


@dataclasses.dataclass(eq=False)
class BackIlluminatedSiliconSensorMaterial(_materials.AbstractMaterial):
    """The light-sensitive material of a back-illuminated silicon sensor
    (``optika/sensors/materials/_materials.py:1723-2282``), by the method of
    Stern 1994.  Thicknesses and roughnesses in nm.  In a trace its
    efficiency -- the s/p average of the Beer-Lambert absorbance in the
    silicon (``:2231-2243``) -- is evaluated per ray inside the trace
    kernels (the oxide over the substrate is a multilayer stack of the
    system); index and attenuation pass through and it is not a mirror.
    ``depletion`` and ``eqe_measured`` are kept for the interface; the
    charge-diffusion and measured-EQE paths are not built."""

    thickness_oxide: float = 0.0
    thickness_implant: float = 0.0
    thickness_substrate: float = 0.0
    roughness_oxide: float = 0.0
    roughness_substrate: float = 0.0
    cce_backsurface: float = 0.0
    depletion: Any = None
    eqe_measured: Any = None
    chemical: str = "Si"
    chemical_oxide: str = "SiO2"

    def __post_init__(self):
        self.kind = _materials.MATERIAL_SILICON_SENSOR

    def _layers(self):
        """(oxide layer, substrate layer) of the multilayer stack."""
        oxide = _materials.Layer(chemical=self.chemical_oxide, thickness=float(self.thickness_oxide),
                                 interface=_materials.profiles.ErfInterfaceProfile(
                                     width=float(self.roughness_oxide)))
        substrate = _materials.Layer(chemical=self.chemical,
                                     thickness=float(self.thickness_substrate),
                                     interface=_materials.profiles.ErfInterfaceProfile(
                                         width=float(self.roughness_substrate)))
        return oxide, substrate

    def transmittance(self, wavelength, direction=None, n=1, normal=None):
        """``_materials.py:1791-1830``."""
        return transmittance(wavelength, _cos_incidence(direction, normal), n,
                             self.thickness_oxide, self.thickness_substrate, self.chemical_oxide,
                             self.chemical, self.roughness_oxide, self.roughness_substrate)

    def absorbance(self, wavelength, direction=None, n=1, normal=None):
        """``_materials.py:1832-1871``."""
        return absorbance(wavelength, _cos_incidence(direction, normal), n, self.thickness_oxide,
                          self.thickness_substrate, self.chemical_oxide, self.chemical,
                          self.roughness_oxide, self.roughness_substrate)

    def charge_collection_efficiency(self, wavelength, direction=None, n=1, normal=None):
        """``_materials.py:1873-1921``."""
        from . import _ops
        from .chemicals import Chemical

        n_sub = Chemical(self.chemical).n(_np(wavelength).astype(np.float64))
        d_sub = _ops.snells_law_scalar(_cos_incidence(direction, normal), n, n_sub)
        return charge_collection_efficiency(absorption_effective(wavelength, n_sub, d_sub),
                                            self.thickness_implant, self.cce_backsurface)

    def quantum_efficiency_effective(self, wavelength, direction=None, n=1, normal=None):
        """``_materials.py:1923-1966``."""
        return quantum_efficiency_effective(
            wavelength, _cos_incidence(direction, normal), n, self.thickness_oxide,
            self.thickness_implant, self.thickness_substrate, self.cce_backsurface,
            self.chemical_oxide, self.chemical, self.roughness_oxide, self.roughness_substrate)
