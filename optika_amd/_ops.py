"""Per-operator device evaluation (optk_eval).

The reference exposes every operator on its own -- ``sag(position)``,
``sag.normal``, ``sag.intercept``, ``aperture(position)``,
``rulings.incident_effective``, ``material.efficiency``, ``snells_law``,
``multilayer_efficiency`` ... (``optika/sags/_abc.py:47-122``,
``optika/apertures/_apertures.py:68-102``, ``optika/rulings/_rulings.py:24-221``,
``optika/materials/_materials.py:28-78``).  Each function here builds a
one-surface system around the operator and runs the same device code the
fused trace inlines, through ``optk_eval``.  Inputs broadcast like numpy.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, _trace
from . import materials as _mat
from . import rulings as _rl
from .rays import RayVectorArray
from .surfaces import Surface
from .vectors import Cartesian3dVectorArray, PolarizationVectorArray, as_vector3

__all__ = [
    "sag_value", "sag_normal", "sag_intercept", "sag_propagate",
    "aperture_mask", "aperture_clip_rays",
    "ruling_spacing", "rulings_incident_effective", "rulings_efficiency", "incident_effective_fn",
    "material_index_refraction", "material_attenuation", "material_efficiency",
    "snells_law", "snells_law_scalar", "multilayer_coefficients", "multilayer_efficiency",
]


def _t(v, dev) -> torch.Tensor:
    if isinstance(v, torch.Tensor):
        return v.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)


def _broadcast(vals, dev):
    ts = torch.broadcast_tensors(*[_t(v, dev) for v in vals])
    shape = ts[0].shape
    return [t.reshape(-1).contiguous() for t in ts], shape


def _run(surface_like, op, ins, nout, iparam=0, mask=False, stack=0, system=None):
    dev = _trace.device()
    flat, shape = _broadcast(ins, dev)
    n = flat[0].numel()
    cs = system if system is not None else _trace.compiled([surface_like])
    outs = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(nout)]
    m = torch.empty(n, dtype=torch.uint8, device=dev) if mask else None
    a = _lib.EvalArgs()
    for k, t in enumerate(flat):
        a.in_[k] = t.data_ptr()
    for k, t in enumerate(outs):
        a.out[k] = t.data_ptr()
    a.out_mask = m.data_ptr() if m is not None else None
    a.surface = 0
    a.iparam = int(iparam)
    a.stack = int(stack)
    _lib.check(_lib.lib().optk_eval(cs.handle, op, ctypes.byref(a), n, _trace.stream_handle()))
    outs = [o.view(shape) for o in outs]
    if mask:
        return m.view(shape).bool(), outs
    return outs


# -- sags -------------------------------------------------------------------


def sag_value(sag, position):
    p = as_vector3(position)
    (z,) = _run(Surface(sag=sag), _lib.OP_SAG_VALUE, [p.x, p.y, p.z], 1)
    return z


def sag_normal(sag, position):
    p = as_vector3(position)
    x, y, z = _run(Surface(sag=sag), _lib.OP_SAG_NORMAL, [p.x, p.y, p.z], 3)
    return Cartesian3dVectorArray(x, y, z)


def _sag_op(sag, rays: RayVectorArray):
    p, d = rays.position, rays.direction
    px, py, pz, dx, dy, dz, inten = _run(
        Surface(sag=sag), _lib.OP_SAG_INTERCEPT,
        [p.x, p.y, p.z, d.x, d.y, d.z, rays.attenuation, rays.intensity], 7)
    return Cartesian3dVectorArray(px, py, pz), Cartesian3dVectorArray(dx, dy, dz), inten


def sag_intercept(sag, rays: RayVectorArray) -> RayVectorArray:
    """``AbstractSag.intercept``: new positions (and, for a rotated sag frame,
    the round-tripped directions), intensity unchanged."""
    pos, d, _ = _sag_op(sag, rays)
    return rays.replace(position=pos, direction=d)


def sag_propagate(sag, rays: RayVectorArray) -> RayVectorArray:
    """``AbstractSag.propagate_rays`` (``optika/sags/_abc.py:109-122``)."""
    pos, d, inten = _sag_op(sag, rays)
    return rays.replace(position=pos, direction=d, intensity=inten)


# -- apertures ----------------------------------------------------------------


def aperture_mask(aperture, position):
    p = as_vector3(position)
    m, _ = _run(Surface(aperture=aperture), _lib.OP_APERTURE, [p.x, p.y, p.z], 0, mask=True)
    return m


def aperture_clip_rays(aperture, rays: RayVectorArray) -> RayVectorArray:
    """``AbstractAperture.clip_rays`` (``_apertures.py:82-102``)."""
    target = rays.direction if aperture.clips_direction else rays.position
    m = aperture_mask(aperture, target)
    u = rays.unvignetted
    if isinstance(u, torch.Tensor):
        u = u.to(m.device) & m
    else:
        u = m if bool(u) else torch.zeros_like(m)
    return rays.replace(unvignetted=u)


# -- rulings ------------------------------------------------------------------


def _rulings_surface(rulings):
    return Surface(rulings=rulings)


def ruling_spacing(spacing, position, normal):
    p, nrm = as_vector3(position), as_vector3(normal)
    s = Surface(rulings=_rl.Rulings(spacing=spacing, diffraction_order=1))
    x, y, z = _run(s, _lib.OP_SPACING, [p.x, p.y, p.z, nrm.x, nrm.y, nrm.z], 3)
    return Cartesian3dVectorArray(x, y, z)


def _paths(rulings):
    return list(range(len(rulings.orders))) if len(rulings.orders) > 1 else [0]


def rulings_incident_effective(rulings, rays: RayVectorArray, normal) -> RayVectorArray:
    """``AbstractRulings.incident_effective`` (``_rulings.py:170-204``); several
    diffraction orders add a leading order axis."""
    nrm = as_vector3(normal)
    kappa = ruling_spacing(rulings.spacing_, rays.position, nrm)
    s = _rulings_surface(rulings)
    d = rays.direction
    res = []
    for path in _paths(rulings):
        res.append(_run(s, _lib.OP_INCIDENT,
                        [rays.wavelength, d.x, d.y, d.z, rays.index_refraction,
                         nrm.x, nrm.y, nrm.z, kappa.x, kappa.y, kappa.z], 3, iparam=path))
    if len(res) == 1:
        x, y, z = res[0]
    else:
        x, y, z = (torch.stack([r[k] for r in res]) for k in range(3))
    return rays.replace(direction=Cartesian3dVectorArray(x, y, z))


def rulings_efficiency(rulings, rays: RayVectorArray, normal):
    nrm = as_vector3(normal)
    s = _rulings_surface(rulings)
    d, p = rays.direction, rays.position
    res = [
        _run(s, _lib.OP_RULINGS_EFF,
             [rays.wavelength, d.x, d.y, d.z, p.x, p.y, p.z, nrm.x, nrm.y, nrm.z], 1,
             iparam=path)[0]
        for path in _paths(rulings)
    ]
    return res[0] if len(res) == 1 else torch.stack(res)


def incident_effective_fn(wavelength, direction, index_refraction, normal, diffraction_order,
                          spacing_rulings, normal_rulings):
    """Free function ``optika.rulings.incident_effective`` (``_rulings.py:24-128``);
    ``wavelength`` and ``spacing_rulings`` share one length unit."""
    a, nrm, g = as_vector3(direction), as_vector3(normal), as_vector3(normal_rulings)
    d = _t(spacing_rulings, _trace.device())
    # the ruling vector is passed in directly; the wavelength is already in
    # the spacing's unit, so the unit scale is one
    sp = _rl.ConstantRulingSpacing(constant=1.0, normal=(1.0, 0.0, 0.0))
    sp.wavelength_scale = 1.0
    s = Surface(rulings=_rl.Rulings(spacing=sp, diffraction_order=int(diffraction_order)))
    x, y, z = _run(s, _lib.OP_INCIDENT,
                   [wavelength, a.x, a.y, a.z, index_refraction, nrm.x, nrm.y, nrm.z,
                    d * _t(g.x, d.device), d * _t(g.y, d.device), d * _t(g.z, d.device)], 3)
    return Cartesian3dVectorArray(x, y, z)


# -- materials ------------------------------------------------------------------


def material_index_refraction(material, rays: RayVectorArray):
    (n,) = _run(Surface(material=material), _lib.OP_INDEX,
                [rays.wavelength, rays.index_refraction], 1)
    return n


def material_attenuation(material, rays: RayVectorArray):
    if isinstance(material, (_mat.Vacuum, _mat.Glass)):
        return 0.0
    return rays.attenuation


def material_efficiency(material, rays: RayVectorArray, normal):
    nrm = as_vector3(normal)
    d = rays.direction
    (e,) = _run(Surface(material=material), _lib.OP_MATERIAL_EFF,
                [rays.wavelength, d.x, d.y, d.z, rays.index_refraction, rays.attenuation,
                 nrm.x, nrm.y, nrm.z], 1)
    return e


def snells_law(direction, index_refraction, index_refraction_new, normal=None,
               is_mirror: bool = False):
    a = as_vector3(direction)
    nrm = Cartesian3dVectorArray(0.0, 0.0, -1.0) if normal is None else as_vector3(normal)
    x, y, z = _run(Surface(), _lib.OP_SNELL,
                   [a.x, a.y, a.z, index_refraction, index_refraction_new,
                    nrm.x, nrm.y, nrm.z], 3, iparam=1 if is_mirror else 0)
    return Cartesian3dVectorArray(x, y, z)


def _split_complex(v, dev):
    # numpy keeps Python scalars in float64 / complex128 (torch would make
    # them float32 / complex64)
    t = torch.as_tensor(np.asarray(v)) if not isinstance(v, torch.Tensor) else v
    if torch.is_complex(t):
        return t.real.to(dev, torch.float64), t.imag.to(dev, torch.float64)
    t = t.to(dev, torch.float64)
    return t, torch.zeros_like(t)


def snells_law_scalar(cos_incidence, index_refraction, index_refraction_new):
    dev = _trace.device()
    c = _split_complex(cos_incidence, dev)
    n1 = _split_complex(index_refraction, dev)
    n2 = _split_complex(index_refraction_new, dev)
    re, im = _run(Surface(), _lib.OP_SNELL_SCALAR, [*c, *n1, *n2], 2)
    return torch.complex(re, im)


def _multilayer(wavelength, direction, n, layers, substrate):
    dev = _trace.device()
    cre, cim = _split_complex(direction, dev)
    nre, nim = _split_complex(n, dev)
    s = Surface(material=_mat.MultilayerMirror(layers=layers, substrate=substrate))
    return _run(s, _lib.OP_MULTILAYER, [wavelength, cre, cim, nre, nim], 12, stack=0)


def multilayer_coefficients(wavelength, direction=1, n=1, layers=None, substrate=None):
    o = _multilayer(wavelength, direction, n, layers, substrate)
    r = PolarizationVectorArray(torch.complex(o[4], o[5]), torch.complex(o[6], o[7]))
    t = PolarizationVectorArray(torch.complex(o[8], o[9]), torch.complex(o[10], o[11]))
    return r, t


def multilayer_efficiency(wavelength, direction=1, n=1, layers=None, substrate=None):
    o = _multilayer(wavelength, direction, n, layers, substrate)
    return PolarizationVectorArray(o[0], o[1]), PolarizationVectorArray(o[2], o[3])
