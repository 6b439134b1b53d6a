"""Batched traces of many instances of one system layout (tolerance
analysis), each instance with its own number of rays, in one persistent
launch (``optk_batch_*``).

The reference expresses such ensembles as n-dimensional configurations:
surface parameters carry extra named axes and broadcast against the ray
arrays (``SequentialSystem.shape``, optika/systems/_sequential.py:2125-2132;
``Surface.shape``, optika/surfaces.py:386-395), which forces every instance
onto the same dense ray grid.  Here each instance owns a contiguous range of
one flat ray array, and the device reduces per-instance statistics of the
final positions: unvignetted count, throughput, centroid sums and the RMS
spot radius.
"""

from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _compile, _lib, _trace

__all__ = ["BatchedSystem", "STAT_FIELDS"]

STAT_FIELDS = ("count", "intensity_sum", "x_sum", "y_sum", "x2_sum", "y2_sum", "spot_rms",
               "throughput")


def _records(system, sensor_frame: bool):
    obj_t = system.object.transformation if system.object is not None else None
    post = system.sensor.transformation if (sensor_frame and system.sensor is not None) else None
    return _compile.build_records(system.surfaces_all, pre=(obj_t, system.transformation),
                                  pre_is_inverse=(0, 1), post=post)


class BatchedSystem:
    """Instances of one :class:`~optika_amd.systems.SequentialSystem` layout
    (same surface kinds and coatings, different parameters), uploaded once.

    ``trace`` runs ``rayfunction`` of every instance over its own slice of
    the input rays and returns per-(instance, path) statistics, optionally
    with the rays themselves.
    """

    def __init__(self, systems: Sequence, sensor_frame: bool = True):
        systems = list(systems)
        if not systems:
            raise ValueError("need at least one system")
        self.dev = _trace.device()
        self.sensor_frame = sensor_frame
        recs = [_records(s, sensor_frame) for s in systems]
        r0 = recs[0]
        for i, r in enumerate(recs[1:], 1):
            same = (r.num_surfaces == r0.num_surfaces and r.num_paths == r0.num_paths
                    and bytes(r.layers) == bytes(r0.layers) and bytes(r.stacks) == bytes(r0.stacks)
                    and bytes(r.chemicals) == bytes(r0.chemicals)
                    and np.array_equal(r.tables, r0.tables)
                    and r.pre_is_inverse == r0.pre_is_inverse)
            if not same:
                raise ValueError(f"instance {i} does not share the layout of instance 0")
        self.num_instances = len(recs)
        self.num_surfaces = r0.num_surfaces
        self.num_paths = r0.num_paths
        self.object_at_infinity = systems[0].object_is_at_infinity
        S = r0.num_surfaces
        surfaces = (_lib.Surface * max(self.num_instances * S, 1))()
        size = ctypes.sizeof(_lib.Surface)
        for i, r in enumerate(recs):
            if S:
                ctypes.memmove(ctypes.addressof(surfaces) + i * S * size, r.surfaces, S * size)
        pre = (_lib.Affine * (2 * self.num_instances))()
        post = (_lib.Affine * self.num_instances)()
        for i, r in enumerate(recs):
            pre[2 * i], pre[2 * i + 1] = r.pre[0], r.pre[1]
            post[i] = r.post
        base, keep = _compile.system_desc(r0)
        desc = _lib.BatchDesc()
        desc.base = ctypes.pointer(base)
        desc.surfaces = ctypes.cast(surfaces, ctypes.POINTER(_lib.Surface))
        desc.pre = ctypes.cast(pre, ctypes.POINTER(_lib.Affine))
        desc.post = ctypes.cast(post, ctypes.POINTER(_lib.Affine))
        desc.num_instances = self.num_instances
        handle = ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.optk_batch_create(ctypes.byref(desc), ctypes.byref(handle)))
        self.handle = handle
        self._keep = (base, keep, surfaces, pre, post, desc)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().optk_batch_destroy(h)
            except Exception:
                pass
            self.handle = None

    # -- helpers ---------------------------------------------------------
    def _grid(self, field_x, field_y, pupil_x, pupil_y, wavelength, dtype):
        g = _lib.Grid()
        ins = []
        for name, t in (("fx", field_x), ("fy", field_y), ("px", pupil_x), ("py", pupil_y),
                        ("wl", wavelength)):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype
                    and t.is_contiguous() and t.ndim == 1):
                raise ValueError(f"{name}: expected a contiguous 1-D {dtype} device tensor")
            ins.append(t)
        n = ins[0].numel()
        if any(t.numel() != n for t in ins):
            raise ValueError("input arrays differ in length")
        g.field[0], g.field[1] = ins[0].data_ptr(), ins[1].data_ptr()
        g.pupil[0], g.pupil[1] = ins[2].data_ptr(), ins[3].data_ptr()
        g.wavelength = ins[4].data_ptr()
        g.intensity_default = 1.0
        g.object_at_infinity = 1 if self.object_at_infinity else 0
        return g, n

    def _offsets(self, counts, n):
        counts = np.asarray(counts, dtype=np.int64).reshape(-1)
        if counts.size != self.num_instances or (counts < 0).any():
            raise ValueError("counts: one non-negative ray count per instance")
        offsets = np.zeros(self.num_instances + 1, dtype=np.int64)
        np.cumsum(counts, out=offsets[1:])
        if offsets[-1] != n:
            raise ValueError(f"counts sum to {offsets[-1]}, but the inputs hold {n} rays")
        return offsets

    def stats_tensor(self) -> torch.Tensor:
        return torch.empty((self.num_instances, self.num_paths, _lib.BATCH_STATS),
                           dtype=torch.float64, device=self.dev)

    # -- API -------------------------------------------------------------
    def trace(self, field_x, field_y, pupil_x, pupil_y, wavelength, counts,
              precision: str = "f32", tune: _lib.Tune | None = None, graph: bool = True,
              outputs: bool = False, stats: torch.Tensor | None = None, generic: bool = False):
        """Trace every instance's rays.  Returns the statistics tensor
        (instances, paths, 8) with columns ``STAT_FIELDS``, and the output
        rays (a ``_trace._Out``: path-major, all instances back to back)
        when ``outputs`` is set.  ``generic``: the generic kernels even
        where every instance matches a specialised layout (same results in
        float64; float32 to its rounding)."""
        prec = _trace._precision(precision)
        dtype = torch.float64 if prec == _lib.F64 else torch.float32
        g, n = self._grid(field_x, field_y, pupil_x, pupil_y, wavelength, dtype)
        offsets = self._offsets(counts, n)
        if stats is None:
            stats = self.stats_tensor()
        out = None
        if outputs:
            out = _trace._Out(self.num_paths * n, self.dev, dtype,
                              {"wavelength", "intensity", "attenuation", "index_refraction"})
        flags = ((_lib.TRACE_APPLY_POST if self.sensor_frame else 0) | (_lib.BATCH_GRAPH if graph else 0)
                 | (_lib.TRACE_GENERIC if generic else 0))
        L = _lib.lib()
        _lib.check(L.optk_batch_trace(
            self.handle, prec, ctypes.byref(g), offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            ctypes.byref(out.rec) if out is not None else None, ctypes.c_void_p(stats.data_ptr()),
            ctypes.byref(tune) if tune is not None else None, flags, _trace.stream_handle()))
        return (stats, out) if outputs else stats

    def autotune(self, field_x, field_y, pupil_x, pupil_y, wavelength, counts,
                 precision: str = "f32") -> _lib.Tune:
        """Time every launch shape on these inputs; later ``trace`` calls
        without ``tune`` use the fastest.  Synchronises the device."""
        prec = _trace._precision(precision)
        dtype = torch.float64 if prec == _lib.F64 else torch.float32
        g, n = self._grid(field_x, field_y, pupil_x, pupil_y, wavelength, dtype)
        offsets = self._offsets(counts, n)
        stats = self.stats_tensor()
        best = _lib.Tune()
        flags = _lib.TRACE_APPLY_POST if self.sensor_frame else 0
        _lib.check(_lib.lib().optk_batch_autotune(
            self.handle, prec, ctypes.byref(g), offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            ctypes.c_void_p(stats.data_ptr()), flags, _trace.stream_handle(), ctypes.byref(best)))
        return best
