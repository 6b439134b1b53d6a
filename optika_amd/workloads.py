"""The measured workloads: a Newtonian telescope and an EUV imaging
spectrograph, with their seeded ray grids.

Newtonian (3 counted surfaces): parabolic primary (conic -1, R = 2000 mm,
circular aperture r = 100 mm) at z = 1000 mm, a flat 45-degree fold mirror
(40 x 30 mm) at z = 100 mm and a 2048 x 2048 sensor of 15 um pixels, 1000 mm
of path from the primary.

Spectrograph (4 counted surfaces): an off-axis conic primary (R = 3000 mm,
k = -0.95, 120 mm decenter of the parent conic), a convex toroidal grating
(radii 1000 / 980 mm) with variable line spacing
``d(x) = d0 (1 + c1 x + c2 x^2)``, ``d0 = 1/2400 mm``, orders -1 and +1,
400 mm before prime focus so that it re-images at about x5, a 6-vertex
field stop 50 mm before a 2048 x 1024 sensor.  Both mirrors carry a Mo/Si
multilayer: 40 periods of 6.7 nm, Mo fraction 0.4, on SiO2.  Every
multilayer is lit from the side that gives a positive incidence cosine
``-a . n`` (the reference's convention, ``optika/materials/_multilayers.py:930``).

The grids are generated with numpy from fixed seeds so that the device
trace, the CPU oracle and the CPU baseline see identical inputs.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import apertures, materials, rulings, sags, sensors, surfaces, systems
from .transformations import (
    Cartesian3dRotationX,
    Cartesian3dRotationY,
    Cartesian3dTranslation,
    TransformationList,
)
from .vectors import Cartesian2dVectorArray, ObjectVectorArray

__all__ = [
    "Grid",
    "newtonian",
    "newtonian_grid_c1",
    "newtonian_grid_random",
    "spectrograph",
    "spectrograph_grid",
    "Perturbation",
    "tolerance_batch",
    "NEWTONIAN_SURFACES",
    "SPECTROGRAPH_SURFACES",
]

NEWTONIAN_SURFACES = 3
SPECTROGRAPH_SURFACES = 4


@dataclasses.dataclass
class Grid:
    """Physical-unit input rays: field angles (deg), pupil (mm), wavelength (nm)."""

    field_x: np.ndarray
    field_y: np.ndarray
    pupil_x: np.ndarray
    pupil_y: np.ndarray
    wavelength: np.ndarray

    @property
    def num(self) -> int:
        return int(np.size(self.pupil_x))

    def object_vector(self, device=None) -> ObjectVectorArray:
        import torch

        def t(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)

        return ObjectVectorArray(
            wavelength=t(self.wavelength),
            field=Cartesian2dVectorArray(t(self.field_x), t(self.field_y)),
            pupil=Cartesian2dVectorArray(t(self.pupil_x), t(self.pupil_y)),
        )

    def subset(self, count: int, start: int = 0) -> "Grid":
        s = slice(start, start + count)
        return Grid(self.field_x[s], self.field_y[s], self.pupil_x[s], self.pupil_y[s],
                    self.wavelength[s])


# ---------------------------------------------------------------------------
# Newtonian telescope
# ---------------------------------------------------------------------------

Z_PRIMARY = 1000.0
Z_FOLD = 100.0
X_SENSOR = 100.0


def newtonian() -> systems.SequentialSystem:
    primary = surfaces.Surface(
        name="primary",
        sag=sags.ConicSag(radius=-2000.0, conic=-1.0),
        material=materials.Mirror(),
        aperture=apertures.CircularAperture(100.0),
        is_pupil_stop=True,
        transformation=Cartesian3dTranslation(z=Z_PRIMARY),
    )
    fold = surfaces.Surface(
        name="fold",
        material=materials.Mirror(),
        aperture=apertures.RectangularAperture(half_width=(20.0, 15.0)),
        transformation=TransformationList([
            Cartesian3dRotationY(135.0),
            Cartesian3dTranslation(z=Z_FOLD),
        ]),
    )
    sensor = sensors.ImagingSensor(
        name="sensor",
        width_pixel=15e-3,
        num_pixel=Cartesian2dVectorArray(2048, 2048),
        transformation=TransformationList([
            Cartesian3dRotationY(-90.0),
            Cartesian3dTranslation(x=-X_SENSOR, z=Z_FOLD),
        ]),
        is_field_stop=True,
    )
    return systems.SequentialSystem(surfaces=[primary, fold], sensor=sensor)


def newtonian_grid_c1(seed: int = 0) -> Grid:
    """64 x 64 jittered pupil cells over [-100, 100]^2 mm, 5 field angles
    x in {0, +-0.1, +-0.2} deg, 500 nm (20,480 rays)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    num = 64
    edges = np.linspace(-100.0, 100.0, num + 1)
    width = edges[1] - edges[0]
    jx = rng.uniform(0.0, 1.0, size=(num, num))
    jy = rng.uniform(0.0, 1.0, size=(num, num))
    px = edges[:-1][:, None] + jx * width
    py = edges[:-1][None, :] + jy * width
    fields = np.array([-0.2, -0.1, 0.0, 0.1, 0.2])
    fx = np.repeat(fields, num * num)
    px = np.tile(px.ravel(), len(fields))
    py = np.tile(py.ravel(), len(fields))
    n = px.size
    return Grid(fx, np.zeros(n), px, py, np.full(n, 500.0))


def newtonian_grid_random(n: int, seed: int = 1) -> Grid:
    """Pupil uniform on the r = 100 mm disc, field angles uniform in
    +-0.25 deg, wavelengths uniform in 400-700 nm."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = 100.0 * np.sqrt(rng.uniform(0.0, 1.0, n))
    th = 2 * np.pi * rng.uniform(0.0, 1.0, n)
    fx = rng.uniform(-0.25, 0.25, n)
    fy = rng.uniform(-0.25, 0.25, n)
    wl = rng.uniform(400.0, 700.0, n)
    return Grid(fx, fy, r * np.cos(th), r * np.sin(th), wl)


# ---------------------------------------------------------------------------
# Imaging spectrograph
# ---------------------------------------------------------------------------

Z_SPEC_PRIMARY = 2000.0
DECENTER = 120.0
D0 = 1.0 / 2400.0  # mm
C1 = 1e-4  # 1/mm
C2 = 1e-7  # 1/mm^2
WAVELENGTH_CENTER = 60.5  # nm
GRATING_BEFORE_FOCUS = 400.0
GRATING_TO_SENSOR = 2000.0
STOP_BEFORE_SENSOR = 50.0


def mo_si_multilayer() -> materials.MultilayerMirror:
    period = 6.7
    gamma = 0.4
    return materials.MultilayerMirror(
        layers=materials.PeriodicLayerSequence(
            layers=[
                materials.Layer("Si", thickness=(1 - gamma) * period),
                materials.Layer("Mo", thickness=gamma * period),
            ],
            num_periods=40,
        ),
        substrate=materials.Layer("SiO2"),
    )


def _rot_y(v, angle_deg):
    a = math.radians(angle_deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([c * v[0] + s * v[2], v[1], -s * v[0] + c * v[2]])


def _angle_y(direction) -> float:
    """Rotation about y (deg) that takes +z to ``direction`` (in the x-z plane)."""
    return math.degrees(math.atan2(direction[0], direction[2]))


def _spectrograph_layout():
    """Frames of the grating, field stop and sensor from the chief ray."""
    c = -1.0 / 3000.0
    k = -0.95
    x = DECENTER
    g = math.sqrt(1 - (1 + k) * c * c * x * x)
    n = np.array([c * x / g, 0.0, -1.0])
    n /= np.linalg.norm(n)
    a = np.array([0.0, 0.0, 1.0])
    b = a - 2 * np.dot(a, n) * n  # chief ray after the primary
    sag_z = c * x * x / (1 + math.sqrt(1 - (1 + k) * c * c * x * x))
    p_hit = np.array([0.0, 0.0, Z_SPEC_PRIMARY + sag_z])
    # prime focus of the parent conic: on the chief ray, where it meets the
    # parent axis x = -decenter
    t_focus = (-DECENTER - p_hit[0]) / b[0]
    grating_pos = p_hit + (t_focus - GRATING_BEFORE_FOCUS) * b
    grating_angle = _angle_y(b)
    # diffracted chief ray (order -1 at the centre wavelength) in the
    # grating frame: effective incidence a + m w g / d, then reflection
    w = WAVELENGTH_CENTER * 1e-6
    a_eff = np.array([0.0, 0.0, 1.0]) + (-1.0) * (-1.0) * w / D0 * np.array([1.0, 0.0, 0.0])
    nrm = np.array([0.0, 0.0, -1.0])
    au = a_eff @ nrm
    d = -au + math.sqrt(1 + au * au - a_eff @ a_eff)
    b_local = a_eff + d * nrm
    b_world = _rot_y(b_local, grating_angle)
    sensor_pos = grating_pos + GRATING_TO_SENSOR * b_world
    stop_pos = grating_pos + (GRATING_TO_SENSOR - STOP_BEFORE_SENSOR) * b_world
    sensor_angle = _angle_y(b_world)
    return grating_pos, grating_angle, stop_pos, sensor_pos, sensor_angle


@dataclasses.dataclass
class Perturbation:
    """Errors of one built spectrograph: decenters (mm) and tilts (deg) of
    the primary and the grating, relative radius errors, relative errors of
    the three line-spacing coefficients."""

    primary_decenter: tuple = (0.0, 0.0)
    primary_tilt: tuple = (0.0, 0.0)
    primary_radius: float = 0.0
    grating_decenter: tuple = (0.0, 0.0)
    grating_tilt: tuple = (0.0, 0.0)
    grating_radii: tuple = (0.0, 0.0)
    spacing: tuple = (0.0, 0.0, 0.0)

    @classmethod
    def gaussian(cls, rng: np.random.Generator, decenter=0.05, tilt=0.01, radius=1e-3,
                 spacing=5e-3) -> "Perturbation":
        n = lambda k, s: tuple(float(v) for v in rng.normal(0.0, s, k))  # noqa: E731
        return cls(primary_decenter=n(2, decenter), primary_tilt=n(2, tilt),
                   primary_radius=n(1, radius)[0], grating_decenter=n(2, decenter),
                   grating_tilt=n(2, tilt), grating_radii=n(2, radius), spacing=n(3, spacing))


def spectrograph(perturbation: Perturbation | None = None) -> systems.SequentialSystem:
    """The spectrograph, optionally with the manufacturing and alignment
    errors of ``perturbation`` (the layout itself stays the nominal one)."""
    p = perturbation or Perturbation()
    grating_pos, grating_angle, stop_pos, sensor_pos, sensor_angle = _spectrograph_layout()

    def frame(pos, angle, tilt=(0.0, 0.0), decenter=(0.0, 0.0)):
        steps = []
        if tilt != (0.0, 0.0):
            steps += [Cartesian3dRotationX(tilt[0]), Cartesian3dRotationY(tilt[1])]
        steps += [
            Cartesian3dRotationY(angle),
            Cartesian3dTranslation(x=float(pos[0]) + decenter[0], y=float(pos[1]) + decenter[1],
                                   z=float(pos[2])),
        ]
        return TransformationList(steps)

    if p.primary_tilt != (0.0, 0.0) or p.primary_decenter != (0.0, 0.0):
        primary_frame = frame((0.0, 0.0, Z_SPEC_PRIMARY), 0.0, p.primary_tilt, p.primary_decenter)
    else:
        primary_frame = Cartesian3dTranslation(z=Z_SPEC_PRIMARY)
    primary = surfaces.Surface(
        name="primary",
        sag=sags.ConicSag(radius=-3000.0 * (1.0 + p.primary_radius), conic=-0.95,
                          transformation=Cartesian3dTranslation(x=-DECENTER)),
        material=mo_si_multilayer(),
        aperture=apertures.CircularAperture(45.0),
        is_pupil_stop=True,
        transformation=primary_frame,
    )
    e0, e1, e2 = p.spacing
    grating = surfaces.Surface(
        name="grating",
        sag=sags.ToroidalSag(radius=980.0 * (1.0 + p.grating_radii[0]),
                             radius_of_rotation=1000.0 * (1.0 + p.grating_radii[1])),
        material=mo_si_multilayer(),
        rulings=rulings.Rulings(
            spacing=rulings.Polynomial1dRulingSpacing(
                coefficients={0: D0 * (1.0 + e0), 1: D0 * C1 * (1.0 + e1), 2: D0 * C2 * (1.0 + e2)},
                normal=(1.0, 0.0, 0.0),
            ),
            diffraction_order=[-1, 1],
        ),
        aperture=apertures.RectangularAperture(half_width=(30.0, 30.0)),
        transformation=frame(grating_pos, grating_angle, p.grating_tilt, p.grating_decenter),
    )
    field_stop = surfaces.Surface(
        name="field_stop",
        aperture=apertures.RegularPolygonalAperture(radius=13.0, num_vertices=6),
        is_field_stop=True,
        transformation=frame(stop_pos, sensor_angle),
    )
    sensor = sensors.ImagingSensor(
        name="sensor",
        width_pixel=15e-3,
        num_pixel=Cartesian2dVectorArray(2048, 1024),
        transformation=frame(sensor_pos, sensor_angle),
    )
    return systems.SequentialSystem(surfaces=[primary, grating, field_stop], sensor=sensor)


def spectrograph_grid(n: int, seed: int = 2) -> Grid:
    """Pupil uniform on the r = 40 mm disc, field angles uniform in
    +-0.05 deg, wavelengths uniform in 58-63 nm."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = 40.0 * np.sqrt(rng.uniform(0.0, 1.0, n))
    th = 2 * np.pi * rng.uniform(0.0, 1.0, n)
    fx = rng.uniform(-0.05, 0.05, n)
    fy = rng.uniform(-0.05, 0.05, n)
    wl = rng.uniform(58.0, 63.0, n)
    return Grid(fx, fy, r * np.cos(th), r * np.sin(th), wl)


def tolerance_batch(num_instances: int = 4096, seed: int = 3, min_rays: int = 1000,
                    max_rays: int = 64000, point_source: bool = True):
    """Batched tolerance analysis of the spectrograph: ``num_instances``
    perturbed copies (decenters sigma 50 um, tilts sigma 0.01 deg, radii
    sigma 0.1 %, line-spacing coefficients sigma 0.5 %), each traced with
    its own ray count, log-uniform in [min_rays, max_rays).

    Returns (systems, counts, grid): the grid holds every instance's rays
    back to back (instance i owns rays offsets[i]:offsets[i+1], offsets the
    running sum of counts).  With ``point_source`` every ray comes from the
    on-axis field point at the centre wavelength (pupil uniform on the
    r = 40 mm disc), so each instance's RMS spot radius is its image
    quality; otherwise the rays are drawn like :func:`spectrograph_grid`.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    perts = [Perturbation.gaussian(rng) for _ in range(num_instances)]
    counts = np.floor(np.exp(rng.uniform(math.log(min_rays), math.log(max_rays),
                                         num_instances))).astype(np.int64)
    grid = spectrograph_grid(int(counts.sum()), seed=seed + 1000)
    if point_source:
        grid.field_x = np.zeros_like(grid.field_x)
        grid.field_y = np.zeros_like(grid.field_y)
        grid.wavelength = np.full_like(grid.wavelength, WAVELENGTH_CENTER)
    systems_ = []
    for p in perts:
        s = spectrograph(p)
        s.perturbation = p  # the errors this instance was built with
        systems_.append(s)
    return systems_, counts, grid


# -- the reference's stop-solve test systems ----------------------------------

RADIUS_FIELD_NEWTONIAN = 0.05  # degrees
RADIUS_FIELD_GRAZING = 0.25


def stop_newtonian() -> systems.SequentialSystem:
    """The reference's stop-solve Newtonian (``_sequential_test.py:536-576``):
    angular object aperture (field stop), pupil stop 250 mm downstream of
    the primary."""
    return systems.SequentialSystem(
        object=surfaces.Surface(
            name="source",
            aperture=apertures.CircularAperture(
                radius=np.sin(np.deg2rad(RADIUS_FIELD_NEWTONIAN)), unit="direction"),
            is_field_stop=True,
        ),
        surfaces=[
            surfaces.Surface(
                name="primary",
                sag=sags.SphericalSag(radius=-2000.0),
                material=materials.Mirror(),
                aperture=apertures.CircularAperture(radius=50.0),
                transformation=Cartesian3dTranslation(z=500.0),
            ),
            surfaces.Surface(
                name="aperture",
                aperture=apertures.CircularAperture(radius=10.0),
                transformation=Cartesian3dTranslation(z=250.0),
                is_pupil_stop=True,
            ),
        ],
        sensor=sensors.ImagingSensor(
            name="sensor", width_pixel=15e-3, num_pixel=Cartesian2dVectorArray(128, 128),
            transformation=Cartesian3dTranslation(z=-500.0),
        ),
    )


def stop_grazing() -> systems.SequentialSystem:
    """The reference's grazing-incidence spectrograph
    (``_sequential_test.py:600-660``): the paraboloid is the pupil stop."""
    return systems.SequentialSystem(
        object=surfaces.Surface(
            name="source",
            aperture=apertures.CircularAperture(
                radius=np.sin(np.deg2rad(RADIUS_FIELD_GRAZING)), unit="direction"),
            is_field_stop=True,
        ),
        surfaces=[
            surfaces.Surface(
                name="paraboloid",
                sag=sags.ParabolicSag(focal_length=-2000.0),
                material=materials.Mirror(),
                aperture=apertures.CircularAperture(radius=260.0),
                transformation=Cartesian3dTranslation(z=2500.0),
                is_pupil_stop=True,
            ),
            surfaces.Surface(
                name="grating",
                rulings=rulings.Rulings(spacing=10e-3, diffraction_order=1),
                aperture=apertures.RectangularAperture(half_width=60.0),
                transformation=Cartesian3dTranslation(z=1000.0),
            ),
        ],
        sensor=sensors.ImagingSensor(
            name="sensor", width_pixel=15e-3, num_pixel=Cartesian2dVectorArray(2048, 1024),
            transformation=Cartesian3dTranslation(x=26.0, z=480.0),
        ),
    )
