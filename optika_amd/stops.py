"""The stop solve: rays that graze the edges of both the field stop and the
pupil stop, and the physical extent of the field and pupil they give.

Mirrors ``optika/systems/_sequential.py:357-789``:

* ``solve_stop_rays`` -- ``_calc_rayfunction_stops_only`` (``:396-662``):
  for each (pupil stop, field stop) pair, every point on the first stop's
  outline (``aperture.wire``) is joined to every point on the last stop's
  outline by a ray whose free component (direction for a stop in
  millimetres, position for an angular stop) is found by Newton's method
  with a forward-difference Jacobian;
* ``rayfunction_stops`` -- ``_calc_rayfunction_stops`` (``:664-712``): those
  rays propagated back to the object surface;
* ``field_pupil`` / ``denormalize`` -- ``field_min`` .. ``pupil_max`` and
  ``_denormalize_grid`` (``:724-789``).

The Newton iteration runs on the device in one launch per stop pair
(``optk_stop_newton``): one thread per outline ray iterates its own root,
tracing the residual and both Jacobian columns through the surfaces in
registers; the walk back to the object is one ``optk_trace``.  The host
keeps the outlines and the seeds (float64), as the reference does; they do
not depend on the wavelength, so they are made once per stop pair and kept
under the content keys of the pair's records.  The solved rays stay on the
device through the walk back and come to the host in one copy.

A grating with several diffraction orders on the way (between the stops,
or between the first stop and the object) gives the reference's stop rays
the order axis: ``_ray_error`` broadcasts over the rulings'
``diffraction_order``, so the residual, the Newton iterate and the solved
rays carry it, and so do the field and pupil extents and the denormalized
grid.  Here each order path is solved on its own (the surfaces with every
multi-order grating restricted to that path's order); the stop rays, the
extents and the denormalized grid gain a path axis in front of the outline
axes, and the trace of a grid that carries one runs each path with its own
inputs.

The reference's root finder is ``named_arrays.optimize.root_newton`` with
``named_arrays.jacobian`` (``named_arrays~=2.1``, not vendored): Newton steps
``x <- x - J(x)^-1 f(x)`` until every residual component is below
``max_abs_error``, the Jacobian by forward differences of step ``dx``.
Its iteration count and stopping details are not visible here; the solve
converges to the same roots (residual below ``1e-9`` of the target
outline's extent).
"""

from __future__ import annotations

import collections
import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib, _trace
from .rays import RayVectorArray
from .vectors import Cartesian3dVectorArray

__all__ = ["solve_stop_rays", "rayfunction_stops", "field_pupil", "denormalize", "angles",
           "num_stop_paths", "single_path"]

MAX_ITERATIONS = 100


def angles(dx, dy, dz):
    """Azimuth and elevation (degrees) of direction cosines
    (``optika/_util.py:69-97``)."""
    length = np.sqrt(dx * dx + dy * dy + dz * dz)
    return -np.rad2deg(np.arctan2(dx, dz)), -np.rad2deg(np.arcsin(dy / length))


def _apply(t, p, direction: bool = False, inverse: bool = False):
    """A transformation of points or directions p [3][...] (host float64)."""
    if t is None:
        return p
    if inverse:
        t = t.inverse
    a = t.affine()
    out = np.tensordot(a.matrix, p, axes=1)
    if not direction:
        out = out + a.vector.reshape((3,) + (1,) * (p.ndim - 1))
    return out


def _sag(surface, p):
    """z of the surface's sag at the points p [3][...] (x, y used)."""
    if surface.sag is None:
        return np.zeros(p.shape[1:])
    from .vectors import Cartesian3dVectorArray as V

    # one copy each way
    d = _to_device(p)
    z = surface.sag(V(d[0], d[1], d[2]))
    return z.reshape(-1).cpu().numpy().reshape(p.shape[1:])


def _to_device(p) -> torch.Tensor:
    """Host components p [k][...] as one device copy, [k][n] float64."""
    host = np.ascontiguousarray(np.asarray(p, np.float64).reshape(len(p), -1))
    return torch.as_tensor(host).to(_trace.device())


def anchor_surface(subsystem):
    """The first surface after the subsystem's start with optical power -- a
    mirror, a curved sag or rulings -- else the last
    (``_sequential.py:347-370``)."""
    from .sags import NoSag

    for s in subsystem[1:]:
        if s.material is not None and getattr(s.material, "is_mirror", False):
            return s
        if s.sag is not None and not isinstance(s.sag, NoSag):
            return s
        if s.rulings is not None:
            return s
    return subsystem[-1]


def _propagate_device(surfaces, position, direction, wavelength, iterations):
    """Device rays [3][N] through ``surfaces`` (system frame in and out);
    returns position, direction [3][N] and ``iterations`` on the host, in
    one copy."""
    dev = _trace.device()
    wave = torch.as_tensor(np.ascontiguousarray(wavelength, np.float64)).to(dev)
    rays = RayVectorArray(
        wavelength=wave,
        position=Cartesian3dVectorArray(position[0], position[1], position[2]),
        direction=Cartesian3dVectorArray(direction[0], direction[1], direction[2]),
    )
    out = _trace.trace_sequence(list(surfaces), rays)
    if out.position.x.ndim != 1:
        raise NotImplementedError(
            "the stop solve needs a single ray path between the stops "
            "(one diffraction order per grating)"
        )
    back = torch.cat([torch.stack([out.position.x, out.position.y, out.position.z,
                                   out.direction.x, out.direction.y, out.direction.z]).reshape(-1),
                      iterations.to(torch.float64).reshape(1)]).cpu().numpy()
    n = out.position.x.numel()
    return back[0:3 * n].reshape(3, n), back[3 * n:6 * n].reshape(3, n), back[-1]


def single_path(surfaces, path: int) -> list:
    """The surfaces with every multi-order grating restricted to the order
    of diffraction-order path ``path`` (the order axis is shared by all
    gratings, as the reference's named axis is)."""
    from ._compile import _orders_of

    out = []
    for s in surfaces:
        rl = getattr(s, "rulings", None)
        if rl is not None:
            orders = _orders_of(rl)
            if len(orders) > 1:
                s = dataclasses.replace(s, rulings=dataclasses.replace(
                    rl, diffraction_order=orders[path]))
        out.append(s)
    return out


def num_stop_paths(system) -> int:
    """Length of the order axis the stop rays carry: that of the surfaces up
    to the last stop, which hold every stop pair and the walk back to the
    object."""
    from ._compile import _num_paths

    surfaces = system.surfaces_all
    ip = [i for i, s in enumerate(surfaces) if s.is_pupil_stop]
    ifs = [i for i, s in enumerate(surfaces) if s.is_field_stop]
    if not ip or not ifs:
        return 1
    return _num_paths(surfaces[:max(ip[-1], ifs[-1]) + 1])


def _wire(surface, num) -> np.ndarray:
    ap = surface.aperture
    if ap is None:
        raise ValueError(f"stop surface {surface.name!r} has no aperture")
    w = ap.wire(num)
    return np.stack([np.asarray(w.x, np.float64), np.asarray(w.y, np.float64),
                     np.asarray(w.z, np.float64)]), ap.clips_direction


# The seeds of a stop pair (outlines, starting rays, targets, step sizes)
# do not depend on the wavelength: they are kept per pair under the content
# keys of the records they derive from (the surfaces after the first stop,
# the first stop itself) and the sample counts, so an edited surface
# misses.  Most-recently used first out; a few pairs suffice.
_SEEDS: "collections.OrderedDict" = collections.OrderedDict()


def _pair_seeds(sub, samples_pupil_stop, samples_field_stop):
    """Seeds of the stop pair sub[0] .. sub[-1] at one wavelength, host
    float64 with the outline axes [S1][S2], and the compiled records the
    Newton solve traces through."""
    from ._compile import build_records
    from .surfaces import Surface

    s_first, s_last = sub[0], sub[-1]
    forward = _trace.compiled(list(sub[1:]), post=s_last.transformation)
    if forward.num_paths != 1:
        raise NotImplementedError(
            "the stop solve needs a single ray path between the stops "
            "(one diffraction order per grating)"
        )
    key = (forward._key, build_records([s_first]).key(), samples_pupil_stop, samples_field_stop,
           bool(s_first.is_pupil_stop), bool(s_first.is_field_stop))
    hit = _SEEDS.get(key)
    if hit is not None:
        _SEEDS.move_to_end(key)
        return hit
    g1, first_angular = _wire(s_first, samples_pupil_stop)
    g2, last_angular = _wire(s_last, samples_field_stop)
    S1, S2 = g1.shape[1], g2.shape[1]
    shape = (S1, S2)
    gf = np.broadcast_to(g1[:2, :, None], (2,) + shape)  # first stop outline
    gl = np.broadcast_to(g2[:2, None, :], (2,) + shape)  # last stop outline
    if not first_angular:
        # outline in millimetres: the ray starts on it, free direction
        pos = np.zeros((3,) + shape)
        pos[:2] = gf
        pos[2] = _sag(s_first, pos)
        dirs = np.zeros((3,) + shape)
        dirs[2] = 1.0
        variable = "direction"
    else:
        # angular outline: the direction is fixed, free position
        dirs = np.zeros((3,) + shape)
        dirs[:2] = gf
        dirs[2] = np.sqrt(1 - (gf[0] ** 2 + gf[1] ** 2))
        pos = np.zeros((3,) + shape)
        variable = "position"

    # seed: aim at the target point on the last stop when no powered
    # surface lies between the stops, else at the first powered surface
    anchor = anchor_surface(sub)
    if anchor is s_last and not last_angular:
        aim = np.zeros((3,) + shape)
        aim[:2] = gl
        aim[2] = _sag(s_last, aim)
        aim = _apply(s_last.transformation, aim)
    else:
        aim = np.zeros((3,) + shape)
        aim = _apply(anchor.transformation, aim)
    aim = _apply(s_first.transformation, aim, inverse=True)
    if variable == "direction":
        d = aim - pos
        d = d / np.sqrt((d * d).sum(0))
        flip = np.sign(d[2])
        nz = d[2] != 0
        dirs = np.stack([np.where(nz, flip * d[0], 0.0), np.where(nz, flip * d[1], 0.0),
                         np.where(nz, flip * d[2], 1.0)])
    else:
        t = aim[2] / dirs[2]
        pos = np.stack([aim[0] - dirs[0] * t, aim[1] - dirs[1] * t, np.zeros(shape)])
        pos[2] = _sag(s_first, pos)
    pos = _apply(s_first.transformation, pos)
    dirs = _apply(s_first.transformation, dirs, direction=True)

    scale = max(np.ptp(g2[0]), np.ptp(g2[1]))
    seeds = dict(
        pos=pos, dirs=dirs, gf=np.array(gf), gl=np.array(gl), shape=shape, variable=variable,
        target="direction" if last_angular else "position",
        tol=1e-9 * max(scale, 1.0), dx=1e-6 if variable == "direction" else 1e-6 * max(scale, 1.0),
        forward=forward,
        first=_trace.compiled([Surface(sag=s_first.sag)]) if variable == "position" else None)
    _SEEDS[key] = seeds
    while len(_SEEDS) > 16:
        _SEEDS.popitem(last=False)
    return seeds


def solve_stop_rays(system, wavelength, samples_pupil_stop: int = 21,
                    samples_field_stop: int = 21, surfaces=None, on_device: bool = False) -> dict:
    """Rays on the first stop joining the outlines of the field and pupil
    stops (``_sequential.py:396-662``), in the system frame.

    ``wavelength``: 1-D array (nm).  Returns ``position``, ``direction``
    [3][S1][S2][W] and the outline grids ``field``, ``pupil`` [2][S1][S2][W],
    with S1 the first stop's outline samples (``samples_pupil_stop``) and S2
    the last stop's (``samples_field_stop``).  ``surfaces``: the system's
    surfaces (default ``surfaces_all``), one diffraction-order path of them
    (``single_path``) for a system with several.  ``on_device``: position
    and direction stay on the device as [3][S1 S2 W] tensors, with the
    solve's largest iteration count (``iterations``, a device scalar) for
    the caller to check; otherwise they are checked and come to the host."""
    if surfaces is None:
        surfaces = system.surfaces_all
    ip = [i for i, s in enumerate(surfaces) if s.is_pupil_stop]
    ifs = [i for i, s in enumerate(surfaces) if s.is_field_stop]
    if not ip:
        raise ValueError("pupil not defined")
    if not ifs:
        raise ValueError("field stop not defined")
    wl = np.atleast_1d(np.asarray(wavelength, np.float64))
    result = None
    while ip and ifs:
        i_p, i_f = ip[0], ifs[0]
        first, last = min(i_p, i_f), max(i_p, i_f)
        sub = surfaces[first:last + 1]
        s_first, s_last = sub[0], sub[-1]
        sd = _pair_seeds(sub, samples_pupil_stop, samples_field_stop)
        S1, S2 = sd["shape"]
        W = wl.size
        shape = (S1, S2, W)
        gf = np.broadcast_to(sd["gf"][..., None], (2,) + shape)
        gl = np.broadcast_to(sd["gl"][..., None], (2,) + shape)
        if s_first.is_pupil_stop:
            ip.pop(0)
            pupil_grid, field_grid = gf, gl
        else:
            ifs.pop(0)
            pupil_grid, field_grid = gl, gf
        wave = np.broadcast_to(wl[None, None, :], shape)
        pos = np.broadcast_to(sd["pos"][..., None], (3,) + shape)
        dirs = np.broadcast_to(sd["dirs"][..., None], (3,) + shape)
        variable = sd["variable"]
        x0 = (dirs if variable == "direction" else pos)[:2].reshape(2, -1)
        # the root and its z (sqrt(1 - x^2 - y^2) for a direction, the first
        # surface's sag for a position), on the device
        d, iters = _newton_device(sd, pos, dirs, wave, gl, x0)
        root = d[9:12]
        if variable == "direction":
            p_dev, d_dev = d[0:3], root
        else:
            p_dev, d_dev = root, d[3:6]
        result = dict(field=np.array(field_grid), pupil=np.array(pupil_grid), wavelength=wave,
                      first=first, s_first=s_first, s_last=s_last, shape=shape)
        if on_device:
            result.update(position=p_dev, direction=d_dev, iterations=iters.max())
        else:
            back = torch.cat([p_dev.reshape(-1), d_dev.reshape(-1),
                              iters.max().to(torch.float64).reshape(1)]).cpu().numpy()
            _check_iterations(back[-1], s_first, s_last)
            result.update(position=back[:-1][:p_dev.numel()].reshape((3,) + shape),
                          direction=back[:-1][p_dev.numel():].reshape((3,) + shape))
    return result


def _check_iterations(count, s_first, s_last):
    if count > MAX_ITERATIONS:
        raise ValueError(
            f"Could not solve for the rays connecting the stop surfaces {s_first.name!r} and "
            f"{s_last.name!r}. If the field stop is only partially reachable at a single "
            f"wavelength (e.g. a spectrograph sensor), consider marking the object surface, "
            f"with an angular (dimensionless, sine of the half-angle) aperture, as the field "
            f"stop instead."
        )


def _newton_device(sd, pos, dirs, wave, gl, x0):
    """Newton's method for every outline ray in one device launch
    (``optk_stop_newton``): each thread iterates its own ray, tracing the
    residual and both forward-difference Jacobian columns through the
    surfaces after the first stop into the last stop's frame.  Returns the
    [12][n] device array (rows 0-5 the starting rays; rows 9-11 the root
    and its z) and the per-ray iteration counts (device)."""
    dev = _trace.device()
    forward, first = sd["forward"], sd["first"]
    n = x0.shape[1]
    # rays, targets and the initial roots in one copy; the roots are solved
    # in place (rows 9 and 10), the free component's z at the root goes to
    # row 11
    host = np.empty((12, n))
    host[0:3] = pos.reshape(3, -1)
    host[3:6] = dirs.reshape(3, -1)
    host[6] = wave.reshape(-1)
    host[7:9] = gl.reshape(2, -1)
    host[9:11] = x0
    d = torch.as_tensor(host).to(dev)
    x = d[9:11]
    iters = torch.empty(n, dtype=torch.int32, device=dev)
    a = _lib.StopArgs()
    for k in range(3):
        a.position[k] = d[k].data_ptr()
        a.direction[k] = d[3 + k].data_ptr()
    a.wavelength = d[6].data_ptr()
    a.target[0], a.target[1] = d[7].data_ptr(), d[8].data_ptr()
    a.x[0], a.x[1] = x[0].data_ptr(), x[1].data_ptr()
    a.iterations = iters.data_ptr()
    a.z = d[11].data_ptr()
    a.tol, a.dx = float(sd["tol"]), float(sd["dx"])
    a.free_component = _lib.STOP_FREE_DIRECTION if sd["variable"] == "direction" else _lib.STOP_FREE_POSITION
    a.target_component = (_lib.STOP_TARGET_POSITION if sd["target"] == "position"
                          else _lib.STOP_TARGET_DIRECTION)
    a.max_iter = MAX_ITERATIONS
    _lib.check(_lib.lib().optk_stop_newton(forward.handle, first.handle if first else None,
                                           ctypes.byref(a), n, _trace.stream_handle()))
    return d, iters


def rayfunction_stops(system, wavelength, samples_pupil_stop: int = 21,
                      samples_field_stop: int = 21) -> dict:
    """The stop rays propagated back to the object surface
    (``_sequential.py:664-712``): ``position``, ``direction`` [3][S1][S2][W]
    in the system frame, directions pointing into the system.  With P > 1
    diffraction-order paths up to the last stop (``num_stop_paths``) they
    are [3][P][S1][S2][W], each path solved and walked back on its own, and
    ``num_paths`` = P."""
    P = num_stop_paths(system)
    if P == 1:
        return _rayfunction_stops_path(system, wavelength, samples_pupil_stop, samples_field_stop,
                                       system.surfaces_all)
    per = [_rayfunction_stops_path(system, wavelength, samples_pupil_stop, samples_field_stop,
                                   single_path(system.surfaces_all, p)) for p in range(P)]
    return dict(position=np.stack([r["position"] for r in per], axis=1),
                direction=np.stack([r["direction"] for r in per], axis=1),
                wavelength=per[0]["wavelength"], num_paths=P)


def _rayfunction_stops_path(system, wavelength, samples_pupil_stop, samples_field_stop,
                            surfaces) -> dict:
    """``rayfunction_stops`` over ``surfaces`` (a single order path)."""
    ip = [i for i, s in enumerate(surfaces) if s.is_pupil_stop]
    ifs = [i for i, s in enumerate(surfaces) if s.is_field_stop]
    if not ip:
        raise ValueError("Pupil stop is not defined for this system.")
    if not ifs:
        raise ValueError("Field stop is not defined for this system.")
    index_stop = min(ip[-1], ifs[-1])
    back = surfaces[index_stop::-1]
    r = solve_stop_rays(system, wavelength, samples_pupil_stop, samples_field_stop, surfaces,
                        on_device=True)
    shape = r["shape"]
    # the solved rays walk back to the object without leaving the device;
    # they and the solve's iteration count come to the host in one copy
    p, d, count = _propagate_device(back, r["position"], r["direction"],
                                    r["wavelength"].reshape(-1), r["iterations"])
    _check_iterations(count, r["s_first"], r["s_last"])
    obj = back[-1]
    pl = _apply(obj.transformation, p, inverse=True)
    dl = _apply(obj.transformation, d, direction=True, inverse=True)
    if obj.sag is not None:
        from . import _ops

        q = _to_device(pl)
        nrm = _ops.sag_normal(obj.sag, Cartesian3dVectorArray(q[0], q[1], q[2]))
        nrm = torch.stack([nrm.x, nrm.y, nrm.z]).cpu().numpy()
    else:
        nrm = np.array([0.0, 0.0, -1.0])[:, None]
    away = (dl * nrm).sum(0) > 0
    d = np.where(away[None, :], -d, d)
    pos, dirs = p.reshape((3,) + shape), d.reshape((3,) + shape)
    if system.transformation is not None:
        pos = _apply(system.transformation, pos)
        dirs = _apply(system.transformation, dirs, direction=True)
    return dict(position=pos, direction=dirs, wavelength=r["wavelength"])


def field_pupil(system, rays: dict):
    """(field, pupil) [2][S1][S2][W] (or [2][P][S1][S2][W]) of stop rays:
    angles (degrees) and positions (mm), which is which by where the object
    is (``_sequential.py:724-786``)."""
    p, d = rays["position"], rays["direction"]
    ax, ay = angles(d[0], d[1], d[2])
    ang = np.stack([ax, ay])
    xy = p[:2]
    if system.object_is_at_infinity:
        return ang, xy
    return xy, ang


def _denormalize_device(components, idx, table, dev):
    """``ptp * (c + 1) / 2 + lo`` (``_sequential.py:777-785``) for each
    component c, lo and ptp taken from ``table[c][0 | 1][w]`` at each value's
    wavelength entry w = ``idx`` (None: entry 0).  Each result has the
    broadcast shape of its component and ``idx``; components of one shape
    share one ``optk_denormalize`` launch."""
    ins = []
    for c in components:
        c = c if isinstance(c, torch.Tensor) else torch.as_tensor(np.asarray(c, np.float64))
        ins.append(c.detach().to(device=dev, dtype=torch.float64))
    groups = {}
    for k, c in enumerate(ins):
        shape = tuple(torch.broadcast_shapes(c.shape, idx.shape)) if idx is not None else tuple(c.shape)
        groups.setdefault(shape, []).append(k)
    outs = [None] * len(ins)
    stride = table.shape[1] * table.shape[2]
    for shape, ks in groups.items():
        a = _lib.DenormalizeArgs()
        n = int(np.prod(shape, dtype=np.int64))
        keep = []
        for j, k in enumerate(ks):
            src = ins[k].expand(shape).contiguous()
            dst = torch.empty(shape, dtype=torch.float64, device=dev)
            keep.append(src)
            outs[k] = dst
            a.in_[j], a.out[j] = src.data_ptr(), dst.data_ptr()
            a.lo[j] = table.data_ptr() + 8 * (k * stride)
            a.ptp[j] = table.data_ptr() + 8 * (k * stride + table.shape[2])
        if idx is not None:
            w = idx.expand(shape).contiguous()
            keep.append(w)
            a.table_index = w.data_ptr()
        a.n_table = table.shape[2]
        _lib.check(_lib.lib().optk_denormalize(ctypes.byref(a), n, _trace.stream_handle()))
    return outs


def denormalize(system, wavelength, field, pupil, normalized_field=True, normalized_pupil=True,
                samples: int = 21):
    """``_denormalize_grid`` (``_sequential.py:748-789``): map normalized
    field and pupil coordinates in [-1, 1] onto the extent the stop rays
    span at each ray's wavelength.  ``wavelength``, ``field.x`` ... are
    device tensors (or scalars) of one length; returns the new field and
    pupil components (device float64)."""
    from .vectors import Cartesian2dVectorArray

    dev = _trace.device()
    if isinstance(wavelength, torch.Tensor) and wavelength.ndim > 0:
        # the distinct wavelengths are found on the device; only they (and
        # not a copy of every ray's wavelength) come to the host
        w = wavelength.detach().reshape(-1).to(device=dev, dtype=torch.float64)
        uniq_t, idx = torch.unique(w, sorted=True, return_inverse=True)
        uniq = uniq_t.cpu().numpy()
    else:
        w = wavelength.detach().cpu().numpy() if isinstance(wavelength, torch.Tensor) else wavelength
        wl_host = np.atleast_1d(np.asarray(w, np.float64)).reshape(-1)
        uniq, inverse = np.unique(wl_host, return_inverse=True)
        idx = torch.as_tensor(inverse.reshape(-1), device=dev) if np.ndim(w) > 0 else None
    rays = rayfunction_stops(system, uniq, samples, samples)
    fld, pup = field_pupil(system, rays)

    def extent(v):
        lo = v.min(axis=(-3, -2))
        return lo, v.max(axis=(-3, -2)) - lo  # [2][W], or [2][P][W]

    if not (normalized_field or normalized_pupil):
        return field, pupil
    # (input component, lo, ptp) in the order field x, field y, pupil x, pupil y
    items = []
    if normalized_field:
        lo, ptp = extent(fld)
        items += [(field.x, lo[0], ptp[0]), (field.y, lo[1], ptp[1])]
    if normalized_pupil:
        lo, ptp = extent(pup)
        items += [(pupil.x, lo[0], ptp[0]), (pupil.y, lo[1], ptp[1])]
    P = rays.get("num_paths", 1)
    per_path = []
    for p in range(P):
        pick = (lambda a: a) if P == 1 else (lambda a, p=p: np.asarray(a)[p])
        table = torch.as_tensor(np.stack([np.stack([np.asarray(pick(a), np.float64).reshape(-1),
                                                    np.asarray(pick(b), np.float64).reshape(-1)])
                                          for _, a, b in items]), device=dev)  # [C][2][W]
        per_path.append(_denormalize_device([c for c, _, _ in items], idx, table, dev))
    # one path: the components as they are; several: each component gains
    # the path axis in front ([P][...]), one denormalized grid per path
    outs = per_path[0] if P == 1 else [torch.stack([o[c] for o in per_path])
                                       for c in range(len(items))]
    k = 0
    if normalized_field:
        field = Cartesian2dVectorArray(outs[0], outs[1])
        k = 2
    if normalized_pupil:
        pupil = Cartesian2dVectorArray(outs[k], outs[k + 1])
    return field, pupil
