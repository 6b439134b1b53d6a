"""The two systems the reference's stop-solve tests use
(optika/systems/_sequential_test.py:536-672), built from optika_amd's
configuration classes: a Newtonian telescope whose pupil stop lies
downstream of the primary, and a grazing-incidence spectrograph whose object
(angular aperture) is the field stop.  Lengths in mm, angles in degrees."""

import numpy as np

from optika_amd import apertures, materials, rulings, sags, sensors, surfaces, systems
from optika_amd.transformations import Cartesian3dTranslation
from optika_amd.vectors import Cartesian2dVectorArray

RADIUS_FIELD_NEWTONIAN = 0.05  # degrees
RADIUS_FIELD_GRAZING = 0.25


def newtonian():
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


def grazing():
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
