"""Device dispatch of the fused trace (optk_trace / optk_trace_grid).

PyTorch supplies device memory and the current stream; every ray field is a
contiguous 1-D tensor and goes to the kernel as a raw pointer.  There is no
host fallback: without a GPU and liboptk.so these functions raise.
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict
from typing import Sequence

import numpy as np
import torch

from . import _compile, _lib
from .rays import RayFunctionArray, RayVectorArray
from .vectors import Cartesian3dVectorArray, ObjectVectorArray

__all__ = ["trace_sequence", "trace_system", "device", "compiled"]

# the layouts the trace kernels are specialised for, in the order of the
# library's signature table (optk_system_specialised returns the index),
# and whether their trace depends on the wavelength (an all-mirror system
# without gratings or coatings does not: its specialised kernels never
# read the wavelength array)
SIGNATURE_NAMES = ("newtonian", "spectrograph", "tilted spectrograph")
SIGNATURE_READS_WAVELENGTH = (False, True, True)

_PRECISIONS = {"f64": _lib.F64, "float64": _lib.F64, "f32": _lib.F32, "float32": _lib.F32}

_COMPILED: "OrderedDict[bytes, _compile.CompiledSystem]" = OrderedDict()


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.OptkError(
            "optika_amd traces on the GPU only: no HIP device is visible "
            "(there is no CPU fallback)"
        )
    _lib.lib()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def compiled(surfaces: Sequence, pre=(None, None), pre_is_inverse=(0, 0), post=None):
    """Compile (or reuse) the device records of a surface sequence."""
    rec = _compile.build_records(surfaces, pre, pre_is_inverse, post)
    key = rec.key()
    cs = _COMPILED.get(key)
    if cs is None:
        device()
        cs = _compile.CompiledSystem(rec)
        _COMPILED[key] = cs
        while len(_COMPILED) > 32:
            _COMPILED.popitem(last=False)
    else:
        _COMPILED.move_to_end(key)
    return cs


def _precision(precision: str) -> int:
    try:
        return _PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")


def _as_tensor(v, n, dev, dtype=torch.float64) -> None | torch.Tensor:
    """A contiguous 1-D device tensor of length n, or None for a scalar."""
    if isinstance(v, torch.Tensor) and v.ndim > 0:
        t = v.reshape(-1).to(device=dev, dtype=dtype).contiguous()
    elif isinstance(v, np.ndarray) and v.ndim > 0:
        t = torch.as_tensor(np.ascontiguousarray(v.reshape(-1))).to(device=dev, dtype=dtype)
    else:
        return None
    if t.numel() != n:
        raise ValueError(f"field has {t.numel()} elements, expected {n}")
    return t


def _scalar(v) -> float:
    if isinstance(v, torch.Tensor):
        return float(v.item())
    return float(v)


def _numel(v) -> int | None:
    if isinstance(v, (torch.Tensor, np.ndarray)) and v.ndim > 0:
        return int(np.prod(v.shape))
    return None


def _ptr(t: None | torch.Tensor):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class _Out:
    """Output tensors of one trace and their ``optk_rays`` record."""

    def __init__(self, total: int, dev, dtype, fields: set[str]):
        mk = lambda: torch.empty(total, dtype=dtype, device=dev)  # noqa: E731
        self.t = {}
        self.t["position"] = [mk(), mk(), mk()]
        self.t["direction"] = [mk(), mk(), mk()]
        for f in ("wavelength", "intensity", "attenuation", "index_refraction"):
            if f in fields:
                self.t[f] = mk()
        # torch.bool is one byte holding 0 or 1: the kernel writes it directly
        self.t["unvignetted"] = torch.empty(total, dtype=torch.bool, device=dev)
        r = _lib.Rays()
        for k in range(3):
            r.position[k] = self.t["position"][k].data_ptr()
            r.direction[k] = self.t["direction"][k].data_ptr()
        for f in ("wavelength", "intensity", "attenuation", "index_refraction"):
            if f in self.t:
                setattr(r, f, self.t[f].data_ptr())
        r.unvignetted = self.t["unvignetted"].data_ptr()
        self.rec = r


def _make_rays(out: _Out, shape, scalars: dict) -> RayVectorArray:
    def shp(t):
        return t.view(shape)

    def field(name):
        if name in out.t:
            return shp(out.t[name])
        return scalars[name]

    return RayVectorArray(
        wavelength=field("wavelength"),
        position=Cartesian3dVectorArray(*(shp(t) for t in out.t["position"])),
        direction=Cartesian3dVectorArray(*(shp(t) for t in out.t["direction"])),
        intensity=field("intensity"),
        attenuation=field("attenuation"),
        index_refraction=field("index_refraction"),
        unvignetted=shp(out.t["unvignetted"]),
    )


def _shape(num_paths: int, n: int, num_surfaces: int | None = None):
    shape = (n,) if num_paths == 1 else (num_paths, n)
    if num_surfaces is not None:
        shape = (num_surfaces,) + shape
    return shape


def trace_sequence(surfaces: Sequence, rays: RayVectorArray, accumulate: bool = False,
                   precision: str = "f64", system=None, generic: bool = False) -> RayVectorArray:
    """``propagate_rays`` / ``accumulate_rays`` over ``surfaces`` (or over an
    already compiled ``system`` of them).  ``generic``: the generic kernels
    even where a specialised one exists (same results)."""
    prec = _precision(precision)
    dtype = torch.float64 if prec == _lib.F64 else torch.float32
    dev = device()
    cs = system if system is not None else compiled(surfaces)
    n = rays.num
    P = cs.num_paths

    keep = []
    rin = _lib.Rays()
    for k, c in enumerate(rays.position):
        t = _as_tensor(c, n, dev, dtype)
        if t is None:
            t = torch.full((n,), _scalar(c), dtype=dtype, device=dev)
        keep.append(t)
        rin.position[k] = t.data_ptr()
    for k, c in enumerate(rays.direction):
        t = _as_tensor(c, n, dev, dtype)
        if t is None:
            t = torch.full((n,), _scalar(c), dtype=dtype, device=dev)
        keep.append(t)
        rin.direction[k] = t.data_ptr()
    scalars_in = {}
    for f, dflt in (("wavelength", "wavelength_default"), ("intensity", "intensity_default"),
                    ("attenuation", "attenuation_default"), ("index_refraction", "index_default")):
        v = getattr(rays, f)
        t = _as_tensor(v, n, dev, dtype)
        if t is None:
            setattr(rin, dflt, _scalar(v))
            scalars_in[f] = _scalar(v)
        else:
            keep.append(t)
            setattr(rin, f, t.data_ptr())
    u = rays.unvignetted
    tu = _as_tensor(u, n, dev, torch.uint8)
    if tu is None:
        rin.unvignetted_default = 1 if bool(u if not isinstance(u, torch.Tensor) else u.item()) else 0
    else:
        keep.append(tu)
        rin.unvignetted = tu.data_ptr()

    fields = {"wavelength", "intensity", "attenuation", "index_refraction"}
    out = _Out(P * n, dev, dtype, fields)
    acc = _Out(cs.num_surfaces * P * n, dev, dtype, fields) if accumulate else None
    L = _lib.lib()
    _lib.check(L.optk_trace(cs.handle, prec, ctypes.byref(rin), ctypes.byref(out.rec),
                            ctypes.byref(acc.rec) if acc else None, n,
                            _lib.TRACE_GENERIC if generic else 0, stream_handle()))
    if accumulate:
        return _make_rays(acc, _shape(P, n, cs.num_surfaces), {})
    return _make_rays(out, _shape(P, n), {})


def _grid_fields(grid: ObjectVectorArray, intensity, paths: int = 1):
    """The grid's fields and their common length n.  With ``paths`` > 1 the
    field and pupil components may carry a leading path axis ([paths][n])."""
    vals = {
        "fx": grid.field.x, "fy": grid.field.y,
        "px": grid.pupil.x, "py": grid.pupil.y,
        "wavelength": grid.wavelength,
        "intensity": 1.0 if intensity is None else intensity,
    }

    def per_ray(k, v):
        m = _numel(v)
        if m is not None and paths > 1 and k in ("fx", "fy", "px", "py") and v.ndim == 2:
            if v.shape[0] != paths:
                raise ValueError(f"grid field has {v.shape[0]} paths, expected {paths}")
            return int(v.shape[1])
        return m

    sizes = {s for s in (per_ray(k, v) for k, v in vals.items()) if s is not None}
    if len(sizes) > 1:
        raise ValueError(
            f"grid fields must be scalars or 1-D arrays of one common length, got sizes {sorted(sizes)}"
        )
    n = sizes.pop() if sizes else 1
    return vals, n


def _grid_record(vals, n, dev, dtype, object_at_infinity, keep):
    """The ``optk_grid`` record of the grid fields ``vals`` (1-D device
    tensors of length n or scalars); the tensors it points at go to
    ``keep``.  Returns (record, wavelength tensor or None)."""
    g = _lib.Grid()

    def put(name, arr_field, idx, dflt_field):
        t = _as_tensor(vals[name], n, dev, dtype)
        if t is None:
            if idx is None:
                setattr(g, dflt_field, _scalar(vals[name]))
            else:
                getattr(g, dflt_field)[idx] = _scalar(vals[name])
        else:
            keep.append(t)
            if idx is None:
                setattr(g, arr_field, t.data_ptr())
            else:
                getattr(g, arr_field)[idx] = t.data_ptr()
        return t

    put("fx", "field", 0, "field_default")
    put("fy", "field", 1, "field_default")
    put("px", "pupil", 0, "pupil_default")
    put("py", "pupil", 1, "pupil_default")
    tw = put("wavelength", "wavelength", None, "wavelength_default")
    put("intensity", "intensity", None, "intensity_default")
    g.object_at_infinity = 1 if object_at_infinity else 0
    return g, tw


def _path_grid(grid: ObjectVectorArray, paths: int) -> bool:
    """True when the grid's field or pupil carries a leading diffraction-order
    path axis ([paths][n] tensors, a grid denormalized per path)."""
    comps = (grid.field.x, grid.field.y, grid.pupil.x, grid.pupil.y)
    return paths > 1 and any(isinstance(c, torch.Tensor) and c.ndim == 2 and c.shape[0] == paths
                             for c in comps)


def trace_system(system, grid: ObjectVectorArray, intensity=None, accumulate: bool = False,
                 sensor_frame: bool = False, precision: str = "f64") -> RayFunctionArray:
    """``SequentialSystem.raytrace`` / ``rayfunction`` in one launch.

    A grid whose field or pupil carries a leading diffraction-order path
    axis (normalized coordinates denormalized per path by the stop solve of
    a system with several orders) is traced one path at a time, each path
    through the surfaces restricted to its order, with its own inputs."""
    prec = _precision(precision)
    surfaces = system.surfaces_all
    obj_t = system.object.transformation if system.object is not None else None
    post = system.sensor.transformation if (sensor_frame and system.sensor is not None) else None
    dtype = torch.float64 if prec == _lib.F64 else torch.float32
    dev = device()
    cs = compiled(surfaces, pre=(obj_t, system.transformation), pre_is_inverse=(0, 1), post=post)
    P = cs.num_paths
    per_path = _path_grid(grid, P)
    vals, n = _grid_fields(grid, intensity, P if per_path else 1)

    invariant = _compile.invariant_fields(surfaces)
    fields = {"intensity"} if invariant else {"wavelength", "intensity", "attenuation",
                                               "index_refraction"}
    all_fields = {"wavelength", "intensity", "attenuation", "index_refraction"}
    out = _Out(P * n, dev, dtype, fields)
    flags = _lib.TRACE_APPLY_POST if post is not None else 0
    L = _lib.lib()
    keep = []
    if not per_path:
        g, tw = _grid_record(vals, n, dev, dtype, system.object_is_at_infinity, keep)
        acc = _Out(cs.num_surfaces * P * n, dev, dtype, all_fields) if accumulate else None
        _lib.check(L.optk_trace_grid(cs.handle, prec, ctypes.byref(g), ctypes.byref(out.rec),
                                     ctypes.byref(acc.rec) if acc else None, n, flags,
                                     stream_handle()))
    else:
        from .stops import single_path

        accs = []
        tw = None
        for p in range(P):
            vp = {k: (v[p] if isinstance(v, torch.Tensor) and v.ndim == 2 else v)
                  for k, v in vals.items()}
            g, tw = _grid_record(vp, n, dev, dtype, system.object_is_at_infinity, keep)
            cs_p = compiled(single_path(surfaces, p), pre=(obj_t, system.transformation),
                            pre_is_inverse=(0, 1), post=post)
            # path p of the path-major outputs: elements p * n .. (p + 1) * n
            rec = _path_slice(out, p * n)
            acc_p = _Out(cs_p.num_surfaces * n, dev, dtype, all_fields) if accumulate else None
            _lib.check(L.optk_trace_grid(cs_p.handle, prec, ctypes.byref(g), ctypes.byref(rec),
                                         ctypes.byref(acc_p.rec) if acc_p else None, n, flags,
                                         stream_handle()))
            accs.append(acc_p)
        if accumulate:
            # [surface][path][n], surface-major as the one-launch form
            def merge(ts):
                return torch.stack([t.view(cs.num_surfaces, n) for t in ts], 1).reshape(-1)

            acc = accs[0]
            acc.t = {k: ([merge([a.t[k][c] for a in accs]) for c in range(3)]
                         if isinstance(v, list) else merge([a.t[k] for a in accs]))
                     for k, v in accs[0].t.items()}
    if accumulate:
        outputs = _make_rays(acc, _shape(P, n, cs.num_surfaces), {})
    else:
        wl = _scalar(vals["wavelength"]) if tw is None else (tw if P == 1 else tw.expand(P, n))
        outputs = _make_rays(out, _shape(P, n), {
            "wavelength": wl, "attenuation": 0.0, "index_refraction": 1.0})
    return RayFunctionArray(inputs=grid, outputs=outputs)


def _path_slice(out: "_Out", start: int):
    """An ``optk_rays`` record pointing at elements start.. of every output
    tensor of ``out``."""
    r = _lib.Rays()
    for k in range(3):
        r.position[k] = out.t["position"][k][start:].data_ptr()
        r.direction[k] = out.t["direction"][k][start:].data_ptr()
    for f in ("wavelength", "intensity", "attenuation", "index_refraction"):
        if f in out.t:
            setattr(r, f, out.t[f][start:].data_ptr())
    r.unvignetted = out.t["unvignetted"][start:].data_ptr()
    return r


class GridTracePlan:
    """A prepared ``rayfunction`` / ``raytrace`` launch for a fixed ray count.

    Compiles the system once, preallocates the output tensors and the
    ``optk_grid`` / ``optk_rays`` records, so each call is one C-ABI launch
    on the current stream with no host-side allocation or record building.
    The returned rays are views of the plan's buffers, overwritten by the
    next call.  ``generic``: run the generic kernels even when the system's
    surface kinds have a specialised kernel (same results; for comparisons).
    """

    def __init__(self, system, n: int, sensor_frame: bool = True, precision: str = "f64",
                 generic: bool = False):
        self.prec = _precision(precision)
        self.dtype = torch.float64 if self.prec == _lib.F64 else torch.float32
        self.dev = device()
        surfaces = system.surfaces_all
        obj_t = system.object.transformation if system.object is not None else None
        post = system.sensor.transformation if (sensor_frame and system.sensor is not None) else None
        self.cs = compiled(surfaces, pre=(obj_t, system.transformation), pre_is_inverse=(0, 1),
                           post=post)
        self.n = int(n)
        self.paths = self.cs.num_paths
        self.flags = (_lib.TRACE_APPLY_POST if post is not None else 0) | (
            _lib.TRACE_GENERIC if generic else 0)
        self.invariant = _compile.invariant_fields(surfaces)
        fields = {"intensity"} if self.invariant else {
            "wavelength", "intensity", "attenuation", "index_refraction"}
        self.out = _Out(self.paths * self.n, self.dev, self.dtype, fields)
        self.grid = _lib.Grid()
        self.grid.object_at_infinity = 1 if system.object_is_at_infinity else 0
        self.grid.intensity_default = 1.0
        self.num_surfaces = self.cs.num_surfaces
        self.generic = bool(generic)

    @property
    def reads_wavelength(self) -> bool:
        """Whether a call reads the wavelength array (the generic kernels
        always do)."""
        sig = self.cs.signature
        if sig >= 0 and not self.generic and sig < len(SIGNATURE_READS_WAVELENGTH):
            return SIGNATURE_READS_WAVELENGTH[sig]
        return True

    @property
    def kernel(self) -> str:
        """The name of the kernel a call launches (its template arguments
        abbreviated): the specialised kernel of the system's surface-kind
        signature, or the generic kernel."""
        prec = "f64" if self.prec == _lib.F64 else "f32"
        sig = self.cs.signature
        if sig >= 0 and not self.generic:
            name = SIGNATURE_NAMES[sig] if sig < len(SIGNATURE_NAMES) else str(sig)
            return f"trace_grid_{prec}_spec<{name} surface kinds>"
        return f"trace_grid_{prec} (generic)"

    def __call__(self, field_x, field_y, pupil_x, pupil_y, wavelength, stream=None):
        """Trace one batch; the five inputs are contiguous device tensors of
        the plan's dtype and length."""
        for t in (field_x, field_y, pupil_x, pupil_y, wavelength):
            if t.dtype != self.dtype or t.numel() != self.n or not t.is_contiguous() or not t.is_cuda:
                raise ValueError(f"inputs must be contiguous {self.dtype} device tensors of length {self.n}")
        g = self.grid
        for k, t in enumerate((field_x, field_y)):
            g.field[k] = t.data_ptr()
        for k, t in enumerate((pupil_x, pupil_y)):
            g.pupil[k] = t.data_ptr()
        g.wavelength = wavelength.data_ptr()
        s = stream_handle() if stream is None else ctypes.c_void_p(stream)
        _lib.check(_lib.lib().optk_trace_grid(self.cs.handle, self.prec, ctypes.byref(g),
                                              ctypes.byref(self.out.rec), None, self.n,
                                              self.flags, s))
        scalars = {"wavelength": wavelength if self.paths == 1 else wavelength.expand(self.paths, self.n),
                   "attenuation": 0.0, "index_refraction": 1.0}
        return _make_rays(self.out, _shape(self.paths, self.n), scalars)
