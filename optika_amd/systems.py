"""Sequential optical systems.

``SequentialSystem`` mirrors ``optika.systems.SequentialSystem``
(``optika/systems/_sequential.py:1826-2132``) on the trace path:
``surfaces_all`` (``:92-112``), input-ray generation
(``_calc_rayfunction_input`` ``:791-828``), ``raytrace`` (``:836-925``) and
``rayfunction`` (``:927-988``).

The whole chain -- input-ray generation from (wavelength, field, pupil),
the object and system transformations, every surface, and for
``rayfunction`` the final change into sensor coordinates -- runs as one
fused device kernel launch.

Field and pupil must be given in physical units (``normalized_field=False``,
``normalized_pupil=False``): the stop solve that denormalizes a normalized
grid (``_sequential.py:396-789``) is not part of this build and raises.
"""

from __future__ import annotations

import dataclasses
from typing import Any, Sequence

from .rays import RayFunctionArray
from .surfaces import Surface
from .transformations import AbstractTransformation
from .vectors import ObjectVectorArray

__all__ = ["AbstractSequentialSystem", "SequentialSystem"]


class AbstractSequentialSystem:
    @property
    def object_is_at_infinity(self) -> bool:
        """``_sequential.py:44-63``: no object aperture, or one that clips on
        direction, means the object is at infinity."""
        ap = self.object.aperture if self.object is not None else None
        if ap is None:
            return True
        return ap.clips_direction

    @property
    def surfaces_all(self) -> list:
        """Object + surfaces + sensor; the first surface becomes the field stop
        when none is marked (``_sequential.py:92-112``)."""
        result = [self.object] if self.object is not None else []
        result += list(self.surfaces)
        if self.sensor is not None:
            result.append(self.sensor)
        if not any(s.is_field_stop for s in result):
            result[0] = dataclasses.replace(result[0], is_field_stop=True)
        return result

    def _check_physical(self, normalized_field: bool, normalized_pupil: bool):
        if normalized_field or normalized_pupil:
            raise NotImplementedError(
                "normalized field/pupil coordinates need the stop solve "
                "(optika/systems/_sequential.py:396-789), which this build does "
                "not provide; pass physical coordinates with "
                "normalized_field=False, normalized_pupil=False"
            )

    def raytrace(
        self,
        intensity=None,
        wavelength=None,
        field=None,
        pupil=None,
        axis: None | str = None,
        normalized_field: bool = True,
        normalized_pupil: bool = True,
        accumulate: bool = True,
        precision: str = "f64",
    ) -> RayFunctionArray:
        """Trace input rays through the system, in system coordinates
        (``_sequential.py:836-925``).

        With ``accumulate`` the output tensors carry a leading surface axis
        (object, surfaces..., sensor).
        """
        from . import _trace

        self._check_physical(normalized_field, normalized_pupil)
        grid = self._grid(wavelength, field, pupil)
        return _trace.trace_system(
            self,
            grid,
            intensity=intensity,
            accumulate=accumulate,
            sensor_frame=False,
            precision=precision,
        )

    def rayfunction(
        self,
        intensity=None,
        wavelength=None,
        field=None,
        pupil=None,
        normalized_field: bool = True,
        normalized_pupil: bool = True,
        precision: str = "f64",
    ) -> RayFunctionArray:
        """Rays at the last surface in sensor coordinates
        (``_sequential.py:927-988``)."""
        from . import _trace

        self._check_physical(normalized_field, normalized_pupil)
        grid = self._grid(wavelength, field, pupil)
        return _trace.trace_system(
            self,
            grid,
            intensity=intensity,
            accumulate=False,
            sensor_frame=True,
            precision=precision,
        )

    def _grid(self, wavelength, field, pupil) -> ObjectVectorArray:
        grid = self.grid_input
        if grid is None:
            grid = ObjectVectorArray()
        grid = grid.copy_shallow()
        if wavelength is not None:
            grid.wavelength = wavelength
        if field is not None:
            grid.field = field
        if pupil is not None:
            grid.pupil = pupil
        return ObjectVectorArray(grid.wavelength, grid.field, grid.pupil)


@dataclasses.dataclass(eq=False)
class SequentialSystem(AbstractSequentialSystem):
    surfaces: Sequence[Any] = ()
    object: None | Any = None
    sensor: None | Any = None
    grid_input: None | ObjectVectorArray = None
    axis_surface: str = "surface"
    axis_wavelength: None | str = None
    axis_field: None | tuple = None
    axis_pupil: None | tuple = None
    transformation: None | AbstractTransformation = None
    kwargs_plot: None | dict = None

    def __post_init__(self):
        if self.object is None:
            self.object = Surface()
