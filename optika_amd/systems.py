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

Normalized field and pupil coordinates (the default) are mapped onto the
physical extent found by the stop solve (``stops.py``,
``_sequential.py:396-789``) before the trace; with
``normalized_field=normalized_pupil=False`` they are used as given.
"""

from __future__ import annotations

import dataclasses
import functools
from typing import Any, Sequence

import numpy as np

from .rays import RayFunctionArray
from .surfaces import Surface
from .transformations import AbstractTransformation
from .vectors import Cartesian2dVectorArray, ObjectVectorArray

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

    # -- stops (_sequential.py:293-345, 714-789) -------------------------------
    @property
    def index_field_stop(self) -> int:
        idx = [i for i, s in enumerate(self.surfaces_all) if s.is_field_stop]
        if not idx:
            raise ValueError(
                "Field stop is not defined for this system."
                "Set `is_field_stop=True` for at least one surface in this system."
            )
        return idx[-1]

    @property
    def index_pupil_stop(self) -> int:
        idx = [i for i, s in enumerate(self.surfaces_all) if s.is_pupil_stop]
        if not idx:
            raise ValueError(
                "Pupil stop is not defined for this system."
                "Set `is_pupil_stop=True` for at least one surface in this system."
            )
        return idx[-1]

    @property
    def field_stop(self):
        return self.surfaces_all[self.index_field_stop]

    @property
    def pupil_stop(self):
        return self.surfaces_all[self.index_pupil_stop]

    def _stop_wavelengths(self):
        grid = self.grid_input if self.grid_input is not None else ObjectVectorArray()
        w = grid.wavelength
        if hasattr(w, "detach"):
            w = w.detach().cpu().numpy()
        return np.unique(np.atleast_1d(np.asarray(w, np.float64)))

    @functools.cached_property
    def rayfunction_stops(self) -> dict:
        """Rays on the object surface that graze both stops' outlines
        (21 x 21 samples) at every wavelength of ``grid_input``
        (``_sequential.py:683-696``, cached on first use as there):
        ``position``, ``direction`` [3][21][21][W] and ``wavelength`` (host
        float64)."""
        from . import stops

        return stops.rayfunction_stops(self, self._stop_wavelengths(), 21, 21)

    def _field_pupil(self):
        from . import stops

        return stops.field_pupil(self, self.rayfunction_stops)

    # the extents reduce over both outline axes, the last two before the
    # wavelength axis; a system with several diffraction-order paths up to
    # its last stop keeps the path axis in front: [P][W]
    @property
    def field_min(self) -> Cartesian2dVectorArray:
        """Lower left corner of the field of view, per wavelength of
        ``grid_input`` (degrees at infinity, else mm; ``:724-734``)."""
        f, _ = self._field_pupil()
        return Cartesian2dVectorArray(*f.min(axis=(-3, -2)))

    @property
    def field_max(self) -> Cartesian2dVectorArray:
        f, _ = self._field_pupil()
        return Cartesian2dVectorArray(*f.max(axis=(-3, -2)))

    @property
    def pupil_min(self) -> Cartesian2dVectorArray:
        """Lower left corner of the entrance pupil, per wavelength of
        ``grid_input`` (mm at infinity, else degrees; ``:748-771``)."""
        _, p = self._field_pupil()
        return Cartesian2dVectorArray(*p.min(axis=(-3, -2)))

    @property
    def pupil_max(self) -> Cartesian2dVectorArray:
        _, p = self._field_pupil()
        return Cartesian2dVectorArray(*p.max(axis=(-3, -2)))

    def _physical_grid(self, grid: ObjectVectorArray, normalized_field: bool,
                       normalized_pupil: bool) -> ObjectVectorArray:
        """``_denormalize_grid`` (``_sequential.py:748-789``)."""
        if not normalized_field and not normalized_pupil:
            return grid
        from . import stops

        field, pupil = stops.denormalize(self, grid.wavelength, grid.field, grid.pupil,
                                         normalized_field, normalized_pupil)
        return ObjectVectorArray(grid.wavelength, field, pupil)

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

        grid = self._input_grid(wavelength, field, pupil, normalized_field, normalized_pupil)
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

        grid = self._input_grid(wavelength, field, pupil, normalized_field, normalized_pupil)
        return _trace.trace_system(
            self,
            grid,
            intensity=intensity,
            accumulate=False,
            sensor_frame=True,
            precision=precision,
        )

    # -- pupil reductions (_sequential.py:1208-1512) -----------------------------
    def _grid_defaults(self, wavelength, field, pupil):
        g = self.grid_input if self.grid_input is not None else ObjectVectorArray()
        return (g.wavelength if wavelength is None else wavelength,
                g.field if field is None else field,
                g.pupil if pupil is None else pupil)

    def _pupil_statistics(self, wavelength, field, pupil, normalized_field, normalized_pupil):
        from . import pupil as _pupil

        wavelength, field, pupil = self._grid_defaults(wavelength, field, pupil)
        rays, shape = _pupil.trace_outer(self, wavelength, field, pupil, normalized_field,
                                         normalized_pupil)
        stats = _pupil.reduce(rays.outputs, shape)
        if stats["count"].shape[0] != 1:
            raise NotImplementedError("pupil models of a system with several diffraction orders")
        W, F, P = shape
        fx = rays.inputs.field.x.reshape(W, F, P)[:, :, 0].cpu().numpy()
        fy = rays.inputs.field.y.reshape(W, F, P)[:, :, 0].cpu().numpy()
        wl = rays.inputs.wavelength.reshape(W, F, P)[:, :, 0].cpu().numpy()
        return {k: v[0] for k, v in stats.items()}, wl, fx, fy, P

    def vignetting(self, wavelength=None, field=None, pupil=None, normalized_field: bool = True,
                   normalized_pupil: bool = True, degree: int = 2):
        """Polynomial vignetting model (``_sequential.py:1282-1367``): the
        fraction of unvignetted rays over the pupil at each wavelength and
        field point, normalized to a mean of one over the field, fitted where
        any ray gets through.  ``wavelength`` [W], ``field`` [F] points,
        ``pupil`` [P] points; the W x F x P grid is traced and reduced on the
        device."""
        from .radiometry import PolynomialVignettingModel

        st, wl, fx, fy, P = self._pupil_statistics(wavelength, field, pupil, normalized_field,
                                                   normalized_pupil)
        where = st["count"] > 0
        illumination = st["count"] / P
        mean = np.array([illumination[w][where[w]].mean() if where[w].any() else np.nan
                         for w in range(illumination.shape[0])])
        illumination = illumination / mean[:, None]
        return PolynomialVignettingModel(wavelength=wl, field_x=fx, field_y=fy,
                                         illumination=illumination, degree=degree, where=where)

    def distortion(self, wavelength=None, field=None, pupil=None, normalized_field: bool = True,
                   normalized_pupil: bool = True, degree: int = 2):
        """Polynomial distortion model (``_sequential.py:1208-1280``): the
        mean sensor position over the unvignetted pupil (over every ray
        where none gets through) at each wavelength and field point."""
        from .distortion import PolynomialDistortionModel

        st, wl, fx, fy, P = self._pupil_statistics(wavelength, field, pupil, normalized_field,
                                                   normalized_pupil)
        where = st["count"] > 0
        n = np.where(where, st["count"], P)
        sx = np.where(where, st["sum_x"], st["all_x"]) / n
        sy = np.where(where, st["sum_y"], st["all_y"]) / n
        return PolynomialDistortionModel(wavelength=wl, field_x=fx, field_y=fy, sensor_x=sx,
                                         sensor_y=sy, degree=degree, where=where)

    def area_effective(self, wavelength=None, field=None, pupil=None,
                       normalized_field: bool = True, normalized_pupil: bool = True,
                       seed: int = 0):
        """Effective area versus wavelength (``_sequential.py:1369-1512``).
        ``pupil`` holds the grid's vertex coordinates along x and along y
        (default 11 each over the normalized pupil); rays are traced at one
        random point of every cell with the cell's area as intensity, their
        output intensity summed over the unvignetted rays and averaged over
        the field points.  mm^2 for an object at infinity, else deg^2."""
        from . import pupil as _pupil
        from . import stops
        from .radiometry import InterpolatedEffectiveAreaModel

        wavelength, field, _ = self._grid_defaults(wavelength, field, None)
        wl = np.atleast_1d(np.asarray(wavelength.detach().cpu().numpy() if hasattr(
            wavelength, "detach") else wavelength, np.float64)).reshape(-1)
        if pupil is None:
            vx = vy = np.linspace(-1.0, 1.0, 11)
        else:
            h = lambda v: np.asarray(v.detach().cpu().numpy() if hasattr(v, "detach") else v,  # noqa: E731
                                     np.float64).reshape(-1)
            vx, vy = h(pupil.x), h(pupil.y)
        if normalized_pupil:
            if stops.num_stop_paths(self) > 1:
                raise NotImplementedError("effective area of a system with several diffraction orders")
            uniq, inv = np.unique(wl, return_inverse=True)
            _, pup = stops.field_pupil(self, stops.rayfunction_stops(self, uniq))
            lo, hi = pup.min(axis=(1, 2)), pup.max(axis=(1, 2))
            vxw = (hi[0] - lo[0])[inv][:, None] * (vx + 1) / 2 + lo[0][inv][:, None]
            vyw = (hi[1] - lo[1])[inv][:, None] * (vy + 1) / 2 + lo[1][inv][:, None]
        else:
            vxw = np.broadcast_to(vx, (wl.size, vx.size))
            vyw = np.broadcast_to(vy, (wl.size, vy.size))
        area, cx, cy = _pupil.cell_grid(vxw, vyw, seed)
        rays, shape = _pupil.trace_outer(self, wl, field, Cartesian2dVectorArray(cx, cy),
                                         normalized_field, False, intensity=area)
        st = _pupil.reduce(rays.outputs, shape)
        if st["throughput"].shape[0] != 1:
            raise NotImplementedError("effective area of a system with several diffraction orders")
        return InterpolatedEffectiveAreaModel(wavelength=wl, area=st["throughput"][0].mean(axis=1))

    # -- forward model (_sequential.py:1002-1206) --------------------------------
    def _extents(self, wavelength):
        """Per wavelength: (field lo, field span, pupil lo, pupil span), each
        [2][W], from the stop solve."""
        from . import stops

        if stops.num_stop_paths(self) > 1:
            raise NotImplementedError("image of a system with several diffraction orders")
        uniq, inv = np.unique(wavelength, return_inverse=True)
        fld, pup = stops.field_pupil(self, stops.rayfunction_stops(self, uniq))
        ext = []
        for v in (fld, pup):
            lo, hi = v.min(axis=(1, 2)), v.max(axis=(1, 2))
            ext += [lo[:, inv], (hi - lo)[:, inv]]
        return ext

    def image(self, wavelength, field_x, field_y, radiance, pupil=None,
              normalized_field: bool = False, normalized_pupil: bool = True,
              integrate: bool = True, noise: bool = True, seed: int = 0):
        """Forward model: the spectral radiance of a scene to the electrons
        of every pixel (``_sequential.py:1002-1206``).

        ``wavelength`` (W + 1), ``field_x`` (X + 1) and ``field_y`` (Y + 1)
        are cell vertices, ``radiance`` is [W][X][Y] per cell (photons per
        unit time, wavelength, field measure and pupil measure).
        ``pupil`` holds the pupil grid's vertex coordinates along x and y
        (default one cell over the normalized pupil).

        As ``_rayfunction_from_vertices`` (``:1002-1083``): normalized
        vertices are denormalized at the vertex wavelengths (the stop solve
        at each); each ray carries radiance x cell volume, where the volume
        (``ObjectVectorArray.cell_area``,
        ``optika/vectors/_vectors_object.py:52-133``) is the wavelength
        step times the field cell's measure times the pupil cell's measure:
        an angular field or pupil cell is measured by its solid angle in
        steradians (``solid_angle_cells``), a positional one by its area in
        mm^2; the field measure is averaged over the cell's two wavelength
        vertices, the pupil measure over the cell's wavelength and field
        vertices.  One ray per (wavelength, field, pupil) cell sits at a
        uniformly random point of the cell: every coordinate is the
        multilinear interpolation of the cell's vertices with uniform
        weights (the reference's ``cell_centers(random=True)``; its random
        stream is not reproduced).  The rays are generated and traced on the
        device, binned on the sensor's pixels over the wavelength cells and
        exposed (``measure``).  Returns [W][Nx][Ny] electrons, or [Nx][Ny]
        with ``integrate``."""
        from .pupil import solid_angle_cells

        h = lambda v: np.asarray(v.detach().cpu().numpy() if hasattr(v, "detach") else v,  # noqa: E731
                                 np.float64).reshape(-1)
        wv, fxv, fyv = h(wavelength), h(field_x), h(field_y)
        W, X, Y = wv.size - 1, fxv.size - 1, fyv.size - 1
        R = h(radiance).reshape(W, X, Y)
        pxv, pyv = (np.array([-1.0, 1.0]), np.array([-1.0, 1.0])) if pupil is None else \
            (h(pupil.x), h(pupil.y))
        PX, PY = pxv.size - 1, pyv.size - 1
        # vertex coordinates per wavelength vertex: [W + 1][.. + 1]
        fx_v, fy_v = np.broadcast_to(fxv, (W + 1, X + 1)), np.broadcast_to(fyv, (W + 1, Y + 1))
        px_v, py_v = np.broadcast_to(pxv, (W + 1, PX + 1)), np.broadcast_to(pyv, (W + 1, PY + 1))
        if normalized_field or normalized_pupil:
            flo, fspan, plo, pspan = self._extents(wv)
            if normalized_field:
                fx_v = fspan[0][:, None] * (fxv + 1) / 2 + flo[0][:, None]
                fy_v = fspan[1][:, None] * (fyv + 1) / 2 + flo[1][:, None]
            if normalized_pupil:
                px_v = pspan[0][:, None] * (pxv + 1) / 2 + plo[0][:, None]
                py_v = pspan[1][:, None] * (pyv + 1) / 2 + plo[1][:, None]
        # cell measures (cell_area): the field is an angle for an object at
        # infinity and the pupil a position, else the other way round
        at_inf = self.object_is_at_infinity

        def measure(vx, vy, angular):
            if angular:
                return solid_angle_cells(vx, vy)
            return np.abs(np.diff(vx, axis=-1)[..., :, None] * np.diff(vy, axis=-1)[..., None, :])

        a_field = measure(fx_v, fy_v, at_inf)           # [W + 1][X][Y]
        a_pupil = measure(px_v, py_v, not at_inf)       # [W + 1][PX][PY]
        a_field = (a_field[:-1] + a_field[1:]) / 2      # cell centers along wavelength
        a_pupil = (a_pupil[:-1] + a_pupil[1:]) / 2
        shape = (W, X, Y, PX, PY)
        import torch

        from . import _trace

        # the rays of every cell are generated on the device (a seeded
        # device generator places each ray in its cell)
        dev = _trace.device()
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed))
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev)  # noqa: E731
        u = torch.rand((5,) + shape, generator=gen, device=dev, dtype=torch.float64)
        uw = u[0]
        wvt = T(wv)
        lam = wvt[:-1, None, None, None, None] + (wvt[1:] - wvt[:-1])[:, None, None, None, None] * uw

        def between(v, axis_shape, uc):
            """the coordinate at weight uw between wavelength vertices and
            uc between the cell's two vertices along its own axis"""
            v = T(v)
            lo = v[:-1, :-1].reshape(axis_shape)       # [W][n] at the cell's lower vertex
            hi = v[:-1, 1:].reshape(axis_shape)
            lo2 = v[1:, :-1].reshape(axis_shape)        # next wavelength vertex
            hi2 = v[1:, 1:].reshape(axis_shape)
            a = lo + (lo2 - lo) * uw
            b = hi + (hi2 - hi) * uw
            return a + (b - a) * uc

        fx = between(fx_v, (W, X, 1, 1, 1), u[1])
        fy = between(fy_v, (W, 1, Y, 1, 1), u[2])
        px = between(px_v, (W, 1, 1, PX, 1), u[3])
        py = between(py_v, (W, 1, 1, 1, PY), u[4])
        volume = (T(np.abs(np.diff(wv)))[:, None, None, None, None]
                  * T(a_field)[:, :, :, None, None] * T(a_pupil)[:, None, None, :, :])
        flux = T(R)[:, :, :, None, None] * volume
        flat = lambda a: a.expand(shape).reshape(-1).contiguous()  # noqa: E731
        rays = self.rayfunction(intensity=flat(flux), wavelength=flat(lam),
                                field=Cartesian2dVectorArray(flat(fx), flat(fy)),
                                pupil=Cartesian2dVectorArray(flat(px), flat(py)),
                                normalized_field=False, normalized_pupil=False).outputs
        if rays.position.x.ndim != 1:
            raise NotImplementedError("image of a system with several diffraction orders")
        electrons = self.sensor.measure(rays, wv, noise=noise, seed=seed)
        return electrons.sum(0) if integrate else electrons

    def _input_grid(self, wavelength, field, pupil, normalized_field, normalized_pupil):
        """The physical input grid: with no wavelength, field or pupil given,
        ``grid_input`` read as normalized whatever the flags
        (``_rayfunction_input``, ``_sequential.py:830-834``, ``:888-889``)."""
        if wavelength is None and field is None and pupil is None:
            normalized_field = normalized_pupil = True
        grid = self._grid(wavelength, field, pupil)
        return self._physical_grid(grid, normalized_field, normalized_pupil)

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
