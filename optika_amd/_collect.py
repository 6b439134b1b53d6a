"""Sensor binning on the device (optk_collect).

``AbstractImagingSensor.collect`` (``optika/sensors/_sensors.py:92-171``):
a weighted histogram of rays over ``(wavelength edges, x edges, y edges)``
with ``num_pixel + 1`` linear pixel edges spanning the sensor aperture,
weights ``intensity`` for unvignetted rays, plus the flux-weighted mean
refracted cosine per pixel.  The ideal sensor material does not refract, so
the cosine is ``-d . n = d_z`` for the sensor normal ``(0, 0, -1)``.

The rays are counting-sorted into 32 x 32-pixel tiles and each tile is
summed with float64 LDS atomics by one block (``tiled=False``, or more than
16384 tiles: float64 atomics in HBM); counts and bin assignment are exact,
weighted sums differ from a sequential sum by rounding only.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, _trace
from .vectors import as_vector2

__all__ = ["collect", "histogram", "expose"]


def histogram(rays, wavelength_edges, x_range, y_range, num_x, num_y, where=None, tiled=True):
    """Raw device histogram: returns ``(image, moment)`` float64 tensors of shape
    ``(n_wavelength, num_x, num_y)``.  ``tiled``: counting-sort the rays into
    32 x 32-pixel tiles and sum each tile in LDS (a workspace of ~20 bytes
    per ray from the caching allocator) instead of float64 global atomics."""
    dev = _trace.device()
    n = rays.num
    f32 = isinstance(rays.position.x, torch.Tensor) and rays.position.x.dtype == torch.float32
    dt = torch.float32 if f32 else torch.float64

    def full(v):
        if isinstance(v, torch.Tensor) and v.ndim > 0:
            return v.reshape(-1).to(device=dev, dtype=dt).contiguous()
        return torch.full((n,), float(v), dtype=dt, device=dev)

    px, py = full(rays.position.x), full(rays.position.y)
    dz = full(rays.direction.z)
    wl = full(rays.wavelength)
    inten = full(rays.intensity)
    u = rays.unvignetted
    if isinstance(u, torch.Tensor) and u.ndim > 0:
        mask = u.reshape(-1).to(device=dev, dtype=torch.uint8).contiguous()
    else:
        mask = torch.full((n,), 1 if bool(u) else 0, dtype=torch.uint8, device=dev)
    if where is not None and not (np.ndim(where) == 0 and bool(where)):
        w = torch.as_tensor(where, device=dev).reshape(-1).to(torch.uint8)
        mask = mask & w.expand(n).contiguous()
    if isinstance(wavelength_edges, torch.Tensor):
        wavelength_edges = wavelength_edges.detach().cpu().numpy()
    e_host = np.asarray(wavelength_edges, dtype=np.float64).reshape(-1)
    nw = e_host.size - 1
    # checked on the host: a device check would synchronise every call
    if nw < 1 or not bool(np.all(e_host[1:] > e_host[:-1])):
        raise ValueError("wavelength edges must hold at least two ascending values")
    edges = torch.as_tensor(e_host, device=dev)
    image = torch.zeros((nw, num_x, num_y), dtype=torch.float64, device=dev)
    moment = torch.zeros_like(image)
    a = _lib.CollectArgs()
    a.position[0], a.position[1] = px.data_ptr(), py.data_ptr()
    a.direction_z = dz.data_ptr()
    a.wavelength = wl.data_ptr()
    a.intensity = inten.data_ptr()
    a.unvignetted = mask.data_ptr()
    a.wavelength_edges = edges.data_ptr()
    a.n_wavelength, a.nx, a.ny = nw, int(num_x), int(num_y)
    a.precision = _lib.F32 if f32 else _lib.F64
    a.x_lo, a.x_hi = float(x_range[0]), float(x_range[1])
    a.y_lo, a.y_hi = float(y_range[0]), float(y_range[1])
    a.image, a.moment = image.data_ptr(), moment.data_ptr()
    L = _lib.lib()
    ws = None
    if tiled:
        nbytes = int(L.optk_collect_workspace(ctypes.byref(a), n))
        if nbytes:
            ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            a.workspace, a.workspace_bytes = ws.data_ptr(), nbytes
    _lib.check(L.optk_collect(ctypes.byref(a), n, _trace.stream_handle()))
    return image, moment


def collect(sensor, rays, wavelength, where=True):
    """``(image, cosine)`` of ``AbstractImagingSensor.collect``."""
    w = as_vector2(sensor.width_pixel)
    npx = as_vector2(sensor.num_pixel)
    hx = w.x * npx.x / 2
    hy = w.y * npx.y / 2
    image, moment = histogram(rays, wavelength, (-hx, hx), (-hy, hy), int(npx.x), int(npx.y),
                              where=None if (np.ndim(where) == 0 and bool(where)) else where)
    cosine = torch.where(image > 0, moment / torch.where(image > 0, image, 1.0),
                         torch.ones_like(image))
    return image, cosine


def expose(image, scale: float, read_noise: float = 0.0, noise: bool = True, seed: int = 0):
    """Electrons of an ideal sensor material (``optk_expose``): a Poisson
    draw of ``scale * image`` per pixel plus Gaussian read noise, or the
    expectation without noise.  ``image``: expected photons per unit time
    per pixel (float64 device tensor)."""
    dev = _trace.device()
    img = image if isinstance(image, torch.Tensor) else torch.as_tensor(np.asarray(image))
    img = img.to(device=dev, dtype=torch.float64).contiguous()
    out = torch.empty_like(img)
    a = _lib.ExposeArgs()
    a.photons, a.electrons = img.data_ptr(), out.data_ptr()
    a.scale, a.read_noise = float(scale), float(read_noise)
    a.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    a.noise = 1 if noise else 0
    _lib.check(_lib.lib().optk_expose(ctypes.byref(a), img.numel(), _trace.stream_handle()))
    return out
