"""optika_amd -- an MI355X-native sequential ray tracer with the trace-path
API of sun-data/optika.

Replace ``import optika`` with ``import optika_amd as optika`` on the trace
path: ``systems.SequentialSystem`` (``raytrace``, ``rayfunction``),
``surfaces.Surface.propagate_rays``, ``propagators.propagate_rays`` /
``accumulate_rays``, the ``sags``, ``rulings``, ``apertures`` and
``materials`` operators and ``sensors.ImagingSensor.collect`` keep their
names and arguments.  ``BatchedSystem`` traces many perturbed instances of
one layout, each with its own ray count, in one persistent launch.  Units are fixed instead of carried: millimetres,
nanometres, degrees, 1/mm.

All of them run on the GPU through ``liboptk.so`` (hand-written gfx950 HIP
kernels behind the C ABI of ``include/optk.h``); there is no CPU fallback.
"""

from __future__ import annotations

from . import (
    apertures,
    chemicals,
    materials,
    propagators,
    rays,
    rulings,
    sags,
    sensors,
    surfaces,
    systems,
    transformations,
    vectors,
)
from .batch import STAT_FIELDS, BatchedSystem

__all__ = [
    "BatchedSystem",
    "STAT_FIELDS",
    "apertures",
    "chemicals",
    "materials",
    "propagators",
    "rays",
    "rulings",
    "sags",
    "sensors",
    "surfaces",
    "systems",
    "transformations",
    "vectors",
    "direction",
    "angles",
]


def direction(angles):
    """Direction cosines from (azimuth, elevation) angles in degrees
    (``optika/_util.py:41-73``): ``R_y(phi_x) R_x(phi_y) z``."""
    import torch

    from .vectors import Cartesian3dVectorArray, as_vector2

    a = as_vector2(angles)
    ax = torch.deg2rad(torch.as_tensor(a.x, dtype=torch.float64))
    ay = torch.deg2rad(torch.as_tensor(a.y, dtype=torch.float64))
    return Cartesian3dVectorArray(
        -torch.cos(ay) * torch.sin(ax), -torch.sin(ay), torch.cos(ay) * torch.cos(ax)
    )


def angles(direction):
    """Inverse of :func:`direction` (``optika/_util.py:76-97``), in degrees."""
    import torch

    from .vectors import Cartesian2dVectorArray, as_vector3

    d = as_vector3(direction)
    x = torch.as_tensor(d.x, dtype=torch.float64)
    y = torch.as_tensor(d.y, dtype=torch.float64)
    z = torch.as_tensor(d.z, dtype=torch.float64)
    length = torch.sqrt(x * x + y * y + z * z)
    return Cartesian2dVectorArray(
        -torch.rad2deg(torch.atan2(x, z)), -torch.rad2deg(torch.asin(y / length))
    )
