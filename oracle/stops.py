"""TEST INFRASTRUCTURE ONLY -- numpy float64 restatement of the reference's
stop solve (``optika/systems/_sequential.py:347-789``) on top of the oracle
trace (``oracle/trace.py``).

* ``wire`` -- aperture outlines (``optika/apertures/_apertures.py:338-355``
  circle, ``:548-590`` sector, ``:704-720`` ellipse, ``:794-830`` polygons);
* ``stop_rays`` -- ``_calc_rayfunction_stops_only`` (``:396-662``): rays on
  the first stop joining both stops' outlines, by Newton's method with a
  forward-difference Jacobian (the reference's ``named_arrays`` root finder,
  restated as ``x <- x - J^-1 f`` until every residual component is below
  ``max_abs_error``);
* ``rayfunction_stops`` -- ``_calc_rayfunction_stops`` (``:664-712``);
* ``field_pupil_extent`` -- ``field_min`` .. ``pupil_max`` (``:724-786``);
* ``denormalize`` -- ``_denormalize_grid`` (``:748-789``).

Pinned in tests/test_oracle_stops.py by the reference's own known answers
(``field_max`` of its Newtonian and grazing-incidence test systems equals
the object aperture's half-angle within 1e-6 degrees,
``optika/systems/_sequential_test.py:591-598``, ``:665-672``), by the
paraxial entrance pupil of the Newtonian system, and by forward traces of
the solved rays landing on the stop outlines.  Imported only by tests/.
"""

from __future__ import annotations

import numpy as np

from . import trace as T

__all__ = ["wire", "surfaces_all", "stop_rays", "rayfunction_stops", "field_pupil_extent",
           "denormalize", "angles", "num_paths"]


def angles(d):
    """``optika.angles`` (``optika/_util.py:76-97``): degrees."""
    length = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    return np.stack([-np.degrees(np.arctan2(d[0], d[2])), -np.degrees(np.arcsin(d[1] / length))])


def _segments(points_per_side, nsides, num, side):
    out, total = [], 0
    for k in range(nsides):
        m = int((k + 1) * points_per_side - total)
        total += m
        t = np.linspace(0.0, 1.0, m, endpoint=(total == num))
        out.append(side(k, t))
    return np.concatenate(out, axis=1)


def wire(aperture, num):
    """Outline points [3][num] in surface coordinates."""
    name = type(aperture).__name__
    if name in ("CircularAperture", "EllipticalAperture"):
        az = np.linspace(0.0, 2 * np.pi, num)
        if name == "CircularAperture":
            rx = ry = float(aperture.radius)
        else:
            rx, ry = float(aperture.radius.x), float(aperture.radius.y)
        p = np.stack([rx * np.cos(az), ry * np.sin(az), 0 * az])
    elif name == "CircularSectorAperture":
        r0 = float(aperture.radius)
        a0, a1 = np.radians(float(aperture.angle_start)), np.radians(float(aperture.angle_stop))

        def side(k, t):
            if k == 0:
                r, a = r0 * t, a0 + 0 * t
            elif k == 1:
                r, a = r0 + 0 * t, a0 + (a1 - a0) * t
            else:
                r, a = r0 * (1 - t), a1 + 0 * t
            return np.stack([r * np.cos(a), r * np.sin(a), 0 * t])

        p = _segments(num / 3, 3, num, side)
    else:
        if name == "RectangularAperture":
            hx, hy = float(aperture.half_width.x), float(aperture.half_width.y)
            vx, vy = np.array([hx, -hx, -hx, hx]), np.array([hy, hy, -hy, -hy])
        else:
            vx, vy = T._polygon_vertices(aperture)
        nv = len(vx)
        v = np.stack([vx, vy, np.zeros(nv)])

        def side(k, t):
            a, b = v[:, k], v[:, (k + 1) % nv]
            return a[:, None] + (b - a)[:, None] * t[None, :]

        p = _segments(num / nv, nv, num, side)
    if getattr(aperture, "transformation", None) is not None:
        p = T.transform_apply(aperture.transformation, p)[0]
    return p


def _angular(aperture) -> bool:
    return getattr(aperture, "unit", "mm") == "direction"


def surfaces_all(system):
    """Object, surfaces, sensor; the first becomes the field stop when none
    is marked (``_sequential.py:92-112``): (surface, is_field, is_pupil)."""
    out = ([system.object] if system.object is not None else []) + list(system.surfaces)
    if system.sensor is not None:
        out.append(system.sensor)
    flags = [(s, bool(s.is_field_stop), bool(s.is_pupil_stop)) for s in out]
    if not any(f for _, f, _ in flags):
        s, _, p = flags[0]
        flags[0] = (s, True, p)
    return flags


def _sag_z(surface, p):
    return T.sag_value(T._sag_of(surface), p)


def _anchor(sub):
    for s in sub[1:]:
        m = s.material
        if m is not None and T.material_is_mirror(m):
            return s
        if s.sag is not None and type(s.sag).__name__ != "NoSag":
            return s
        if getattr(s, "rulings", None) is not None:
            return s
    return sub[-1]


def _to_global(t, p, direction=False):
    if t is None:
        return p
    if direction:
        return T.transform_apply(t, np.zeros_like(p), p)[1]
    return T.transform_apply(t, p)[0]


def _to_local(t, p, direction=False):
    if t is None:
        return p
    if direction:
        return T.transform_inverse(t, np.zeros_like(p), p)[1]
    return T.transform_inverse(t, p)[0]


def _trace(surfaces, pos, dirs, wl, path=0):
    """Rays through ``surfaces`` along diffraction-order path ``path`` (a
    sequence without a multi-order surface has the one path 0)."""
    rays = T.make_rays(wl, pos, dirs)
    paths = T.propagate(surfaces, rays)
    r = paths[path if len(paths) > 1 else 0]
    return r["position"], r["direction"]


def num_paths(system) -> int:
    """Length of the diffraction-order axis that the stop rays carry: that of
    the surfaces up to the last stop (the stop pairs and the walk back to the
    object).  A grating with several orders between the stops gives the
    solve's residual, and so its root, the order axis
    (``_sequential.py:396-662``: the named-array broadcast of ``_ray_error``
    over the rulings' ``diffraction_order``); each order is solved on its own."""
    flags = surfaces_all(system)
    last = max(max(i for i, (_, _, p) in enumerate(flags) if p),
               max(i for i, (_, f, _) in enumerate(flags) if f))
    return len(T.system_orders([s for s, _, _ in flags[:last + 1]]))


def stop_rays(system, wavelength, s1=21, s2=21, max_iter=100, path=0):
    """Rays on the first stop (system frame) for the last stop pair, along
    diffraction-order path ``path``:
    dict(position, direction [3][s1][s2][W], field, pupil [2][s1][s2][W])."""
    flags = surfaces_all(system)
    surfaces = [s for s, _, _ in flags]
    ip = [i for i, (_, _, p) in enumerate(flags) if p]
    ifs = [i for i, (_, f, _) in enumerate(flags) if f]
    wl = np.atleast_1d(np.asarray(wavelength, np.float64))
    out = None
    while ip and ifs:
        a, b = min(ip[0], ifs[0]), max(ip[0], ifs[0])
        first_is_pupil = flags[a][2]
        if first_is_pupil:
            ip.pop(0)
        else:
            ifs.pop(0)
        sub = surfaces[a:b + 1]
        first, last = sub[0], sub[-1]
        ap1, ap2 = T._aperture_of(first), T._aperture_of(last)
        w1, w2 = wire(ap1, s1), wire(ap2, s2)
        shape = (w1.shape[1], w2.shape[1], wl.size)
        N = int(np.prod(shape))
        g1 = np.broadcast_to(w1[:2, :, None, None], (2,) + shape).reshape(2, N)
        g2 = np.broadcast_to(w2[:2, None, :, None], (2,) + shape).reshape(2, N)
        lam = np.broadcast_to(wl[None, None, :], shape).reshape(N)
        free_direction = not _angular(ap1)
        if free_direction:
            pos = np.stack([g1[0], g1[1], np.zeros(N)])
            pos[2] = _sag_z(first, pos)
            dirs = np.stack([np.zeros(N), np.zeros(N), np.ones(N)])
        else:
            dirs = np.stack([g1[0], g1[1], np.sqrt(1 - g1[0] ** 2 - g1[1] ** 2)])
            pos = np.zeros((3, N))
        anc = _anchor(sub)
        if anc is last and not _angular(ap2):
            aim = np.stack([g2[0], g2[1], np.zeros(N)])
            aim[2] = _sag_z(last, aim)
            aim = _to_global(last.transformation, aim)
        else:
            aim = _to_global(anc.transformation, np.zeros((3, N)))
        aim = _to_local(first.transformation, aim)
        if free_direction:
            d = aim - pos
            d = d / np.sqrt((d ** 2).sum(0))
            sg = np.sign(d[2])
            ok = d[2] != 0
            dirs = np.where(ok, sg * d, np.array([[0.0], [0.0], [1.0]]))
        else:
            t = aim[2] / dirs[2]
            pos = np.stack([aim[0] - dirs[0] * t, aim[1] - dirs[1] * t, np.zeros(N)])
            pos[2] = _sag_z(first, pos)
        pos = _to_global(first.transformation, pos)
        dirs = _to_global(first.transformation, dirs, direction=True)
        target_direction = _angular(ap2)
        scale = max(np.ptp(w2[0]), np.ptp(w2[1]))
        tol = 1e-9 * max(scale, 1.0)
        step = 1e-6 if free_direction else 1e-6 * max(scale, 1.0)

        def z_of(x):
            if free_direction:
                return np.sqrt(1 - x[0] ** 2 - x[1] ** 2)
            return _sag_z(first, np.stack([x[0], x[1], np.zeros(x.shape[1])]))

        def f(x):
            v = np.stack([x[0], x[1], z_of(x)])
            p, d = (pos, v) if free_direction else (v, dirs)
            p2, d2 = _trace(sub[1:], p, d, lam, path)
            q = _to_local(last.transformation, d2 if target_direction else p2,
                          direction=target_direction)
            return q[:2] - g2

        x = (dirs if free_direction else pos)[:2].copy()
        for _ in range(max_iter):
            f0 = f(x)
            if np.all(np.abs(f0) < tol):
                break
            fa = f(x + np.array([[step], [0.0]]))
            fb = f(x + np.array([[0.0], [step]]))
            J = np.stack([np.stack([(fa[0] - f0[0]) / step, (fb[0] - f0[0]) / step], -1),
                          np.stack([(fa[1] - f0[1]) / step, (fb[1] - f0[1]) / step], -1)], -2)
            x = x - np.linalg.solve(J, f0.T[..., None])[..., 0].T
        else:
            raise ValueError("stop solve did not converge")
        v = np.stack([x[0], x[1], z_of(x)])
        if free_direction:
            dirs = v
        else:
            pos = v
        pupil, field = (g1, g2) if first_is_pupil else (g2, g1)
        out = dict(position=pos.reshape((3,) + shape), direction=dirs.reshape((3,) + shape),
                   pupil=pupil.reshape((2,) + shape), field=field.reshape((2,) + shape),
                   wavelength=lam.reshape(shape))
    return out


def rayfunction_stops(system, wavelength, s1=21, s2=21, path=0):
    """The stop rays walked back to the object (``_sequential.py:664-712``),
    along diffraction-order path ``path``."""
    flags = surfaces_all(system)
    surfaces = [s for s, _, _ in flags]
    ip = [i for i, (_, _, p) in enumerate(flags) if p]
    ifs = [i for i, (_, f, _) in enumerate(flags) if f]
    k = min(ip[-1], ifs[-1])
    back = surfaces[k::-1]
    r = stop_rays(system, wavelength, s1, s2, path=path)
    shape = r["position"].shape[1:]
    p, d = _trace(back, r["position"].reshape(3, -1), r["direction"].reshape(3, -1),
                  r["wavelength"].reshape(-1), path)
    obj = back[-1]
    pl = _to_local(obj.transformation, p)
    dl = _to_local(obj.transformation, d, direction=True)
    nrm = T.sag_normal(T._sag_of(obj), pl)
    flip = (dl * nrm).sum(0) > 0
    d = np.where(flip, -d, d)
    if system.transformation is not None:
        p = _to_global(system.transformation, p)
        d = _to_global(system.transformation, d, direction=True)
    return dict(position=p.reshape((3,) + shape), direction=d.reshape((3,) + shape))


def field_pupil_extent(system, wavelength, s=21, path=0):
    """(field, pupil) as (min [2][W], max [2][W]) pairs, along
    diffraction-order path ``path``."""
    r = rayfunction_stops(system, wavelength, s, s, path=path)
    ang = angles(r["direction"])
    xy = r["position"][:2]
    field, pupil = (ang, xy) if T._object_at_infinity(system) else (xy, ang)
    ext = lambda v: (v.min(axis=(1, 2)), v.max(axis=(1, 2)))  # noqa: E731
    return ext(field), ext(pupil)


def denormalize(system, wavelength, field_x, field_y, pupil_x, pupil_y, s=21, path=0):
    """Normalized coordinates in [-1, 1] onto the stop rays' extent, at one
    wavelength (scalar), along diffraction-order path ``path``."""
    (flo, fhi), (plo, phi) = field_pupil_extent(system, [wavelength], s, path=path)
    fx = (fhi[0, 0] - flo[0, 0]) * (np.asarray(field_x) + 1) / 2 + flo[0, 0]
    fy = (fhi[1, 0] - flo[1, 0]) * (np.asarray(field_y) + 1) / 2 + flo[1, 0]
    px = (phi[0, 0] - plo[0, 0]) * (np.asarray(pupil_x) + 1) / 2 + plo[0, 0]
    py = (phi[1, 0] - plo[1, 0]) * (np.asarray(pupil_y) + 1) / 2 + plo[1, 0]
    return fx, fy, px, py
