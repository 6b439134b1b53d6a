"""Optical surfaces.

``Surface`` mirrors ``optika.surfaces.Surface`` (``optika/surfaces.py:282-395``):
a sag, a material, an aperture, optional rulings and a transformation from
the surface frame into the system frame.  ``propagate_rays`` performs the
per-surface step of ``AbstractSurface.propagate_rays``
(``optika/surfaces.py:123-198``) by tracing through a one-surface system
with the fused device kernel.
"""

from __future__ import annotations

import dataclasses
from typing import Any

from . import materials as _materials
from . import sags as _sags
from .transformations import AbstractTransformation

__all__ = ["AbstractSurface", "Surface"]


class AbstractSurface:
    """A ray propagator made of sag, material, aperture and rulings."""

    @property
    def is_stop(self) -> bool:
        return self.is_field_stop or self.is_pupil_stop

    def propagate_rays(self, rays):
        from . import propagators

        return propagators.propagate_rays([self], rays)


@dataclasses.dataclass(eq=False)
class Surface(AbstractSurface):
    name: None | str = None
    sag: Any = None
    material: Any = None
    aperture: Any = None
    aperture_mechanical: Any = None
    rulings: Any = None
    is_field_stop: bool = False
    is_pupil_stop: bool = False
    transformation: None | AbstractTransformation = None
    kwargs_plot: None | dict = None

    def __post_init__(self):
        if self.sag is None:
            self.sag = _sags.NoSag()
        if self.material is None:
            self.material = _materials.Vacuum()
