"""Reductions over the pupil of traced rays: the data behind
``SequentialSystem.vignetting``, ``distortion`` and ``area_effective``
(``optika/systems/_sequential.py:1208-1512``).

The reference traces a named-axis grid (wavelength x field x pupil) and
reduces along the pupil axes with numpy.  Here the grid is the outer product
of flat arrays -- W wavelengths, F field points, P pupil points, pupil
fastest -- traced in one launch of the fused kernel (``optk_trace_grid``),
and every (path, wavelength, field) segment of P rays is reduced on the
device by one wavefront (``optk_pupil_reduce``: counts, position sums, the
area-weighted throughput; fixed-order sums).  Only the [paths][W][F]
results come back to the host.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, _trace
from .vectors import Cartesian2dVectorArray, ObjectVectorArray

__all__ = ["trace_outer", "reduce", "cell_grid", "solid_angle_cells"]


def _flat(v, dev) -> torch.Tensor:
    if isinstance(v, torch.Tensor):
        return v.detach().reshape(-1).to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.atleast_1d(np.asarray(v, np.float64)).reshape(-1), device=dev)


def trace_outer(system, wavelength, field, pupil, normalized_field=True, normalized_pupil=True,
                intensity=None):
    """Trace every (wavelength, field point, pupil point) combination into
    the sensor frame.  ``pupil`` components (and ``intensity``) are [P] or
    [W][P] (a pupil per wavelength).  Returns the ray function and
    (W, F, P)."""
    dev = _trace.device()
    wl = _flat(wavelength, dev)
    fx, fy = torch.broadcast_tensors(_flat(field.x, dev), _flat(field.y, dev))
    W, F = wl.numel(), fx.numel()

    def per_w(v):
        t = v.detach().to(device=dev, dtype=torch.float64) if isinstance(v, torch.Tensor) else \
            torch.as_tensor(np.asarray(v, np.float64), device=dev)
        return t.reshape(W, -1) if t.ndim == 2 else t.reshape(1, -1)

    px, py = torch.broadcast_tensors(per_w(pupil.x), per_w(pupil.y))
    P = px.shape[-1]
    shape = (W, F, P)

    g = ObjectVectorArray(
        wavelength=wl[:, None, None].expand(shape).reshape(-1).contiguous(),
        field=Cartesian2dVectorArray(fx[None, :, None].expand(shape).reshape(-1).contiguous(),
                                     fy[None, :, None].expand(shape).reshape(-1).contiguous()),
        pupil=Cartesian2dVectorArray(px[:, None, :].expand(shape).reshape(-1).contiguous(),
                                     py[:, None, :].expand(shape).reshape(-1).contiguous()),
    )
    if intensity is not None and np.ndim(intensity) > 0:
        it = per_w(intensity)
        intensity = it[:, None, :].expand(shape).reshape(-1).contiguous()
    g = system._physical_grid(g, normalized_field, normalized_pupil)
    rays = _trace.trace_system(system, g, intensity=intensity, accumulate=False, sensor_frame=True)
    return rays, shape


def reduce(outputs, shape, weight=None) -> dict:
    """Per (path, wavelength, field) segment of a ``trace_outer`` result:
    ``count`` of unvignetted rays, ``sum_x``/``sum_y`` over them,
    ``all_x``/``all_y`` over every ray and ``throughput`` (Σ weight ×
    intensity over the unvignetted rays).  Host arrays [paths][W][F]."""
    W, F, P = shape
    x = outputs.position.x.reshape(-1).contiguous()
    y = outputs.position.y.reshape(-1).contiguous()
    m = outputs.unvignetted.reshape(-1).contiguous()
    paths = x.numel() // (W * F * P)
    inten = outputs.intensity
    if not isinstance(inten, torch.Tensor) or inten.ndim == 0:
        inten = torch.full_like(x, float(inten))
    inten = inten.reshape(-1).contiguous().to(x.dtype)
    segs = paths * W * F
    out = {k: torch.zeros(segs, dtype=torch.float64, device=x.device)
           for k in ("count", "sum_x", "sum_y", "all_x", "all_y", "throughput")}
    if x.dtype != torch.float64:
        x, y, inten = x.double(), y.double(), inten.double()
    w = None if weight is None else _flat(weight, x.device)
    a = _lib.PupilArgs()
    a.x, a.y, a.intensity = x.data_ptr(), y.data_ptr(), inten.data_ptr()
    a.unvignetted = m.data_ptr()
    a.weight = w.data_ptr() if w is not None else None
    a.per_segment = P
    for k, t in out.items():
        setattr(a, k, t.data_ptr())
    _lib.check(_lib.lib().optk_pupil_reduce(ctypes.byref(a), segs, _trace.stream_handle()))
    return {k: t.cpu().numpy().reshape(paths, W, F) for k, t in out.items()}


def cell_grid(vx, vy, seed: int = 0):
    """Cells of the tensor grid of vertices vx [..., Nx] x vy [..., Ny]
    (leading axes broadcast, e.g. one row per wavelength): areas
    |dx dy| and one uniformly random point in each cell (the reference's
    ``cell_centers(random=True)``; its random stream is not reproduced).
    Returns areas, x, y, each [..., (Nx - 1) * (Ny - 1)], x fastest along y."""
    vx, vy = np.asarray(vx, np.float64), np.asarray(vy, np.float64)
    dx = np.diff(vx, axis=-1)
    dy = np.diff(vy, axis=-1)
    rng = np.random.default_rng(seed)
    nx, ny = dx.shape[-1], dy.shape[-1]
    ux = rng.uniform(0, 1, (nx, ny))
    uy = rng.uniform(0, 1, (nx, ny))
    area = np.abs(dx[..., :, None] * dy[..., None, :])
    X = vx[..., :-1, None] + dx[..., :, None] * ux
    Y = vy[..., None, :-1] + dy[..., None, :] * uy
    X, Y = np.broadcast_to(X, area.shape), np.broadcast_to(Y, area.shape)
    lead = area.shape[:-2]
    return (area.reshape(lead + (-1,)), X.reshape(lead + (-1,)), Y.reshape(lead + (-1,)))


def _unit_direction(ax_deg, ay_deg):
    """``optika.direction`` (``optika/_util.py:41-73``) of angles in degrees:
    (-cos y sin x, -sin y, cos y cos x), components on the last axis."""
    x, y = np.broadcast_arrays(np.deg2rad(ax_deg), np.deg2rad(ay_deg))
    return np.stack([-np.cos(y) * np.sin(x), -np.sin(y), np.cos(y) * np.cos(x)], axis=-1)


def _triangle_solid_angle(a, b, c):
    """Signed solid angle of the spherical triangle of unit vectors a, b, c
    (Van Oosterom and Strackee, IEEE Trans. Biomed. Eng. 30, 125 (1983)):
    tan(omega / 2) = a . (b x c) / (1 + a.b + b.c + c.a)."""
    triple = np.sum(a * np.cross(b, c), axis=-1)
    den = 1.0 + np.sum(a * b, axis=-1) + np.sum(b * c, axis=-1) + np.sum(c * a, axis=-1)
    return 2.0 * np.arctan2(triple, den)


def solid_angle_cells(ax, ay):
    """Solid angle (steradians) of every cell of the angle grid with vertices
    ax [..., Nx + 1] x ay [..., Ny + 1] (degrees; leading axes broadcast):
    the cell's four vertex directions (``optika.direction``) bound a
    spherical quadrilateral with great-circle edges, summed as two
    triangles.  The reference weights an angular field (or pupil) cell by
    this measure, ``optika.direction(field).solid_angle_cell(axis)``
    (``optika/vectors/_vectors_object.py:111-118``); for small cells it
    tends to |dx (sin y1 - sin y0)| (x, y in radians), the area of the
    coordinate cell on the unit sphere.  Returns |omega| [..., Nx, Ny]."""
    ax, ay = np.asarray(ax, np.float64), np.asarray(ay, np.float64)
    lead = np.broadcast_shapes(ax.shape[:-1], ay.shape[:-1])
    ax = np.broadcast_to(ax, lead + ax.shape[-1:])
    ay = np.broadcast_to(ay, lead + ay.shape[-1:])
    d = _unit_direction(ax[..., :, None], ay[..., None, :])  # [..., Nx+1, Ny+1, 3]
    v00, v10 = d[..., :-1, :-1, :], d[..., 1:, :-1, :]
    v11, v01 = d[..., 1:, 1:, :], d[..., :-1, 1:, :]
    omega = _triangle_solid_angle(v00, v10, v11) + _triangle_solid_angle(v00, v11, v01)
    return np.abs(omega)
