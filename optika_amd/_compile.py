"""Compile a sequence of surfaces into the flat records of include/optk.h.

The reference walks a Python list of surface objects and dispatches every
operator through Python (``optika/propagators.py:38-39``).  Here the list is
flattened once, on the host, into plain records -- one ``optk_surface`` per
surface, multilayer stacks as blocks of layers, and every table (n/k data,
measured efficiencies, polygon vertices) packed into one float64 array --
and uploaded to the device, where one kernel launch walks all surfaces.

Compiled systems are cached on the host object and reused while their
record bytes are unchanged.
"""

from __future__ import annotations

import ctypes
import dataclasses
from typing import Sequence

import numpy as np

from . import _lib
from . import apertures as _ap
from . import materials as _mat
from . import rulings as _rl
from .transformations import AbstractTransformation

__all__ = ["CompiledSystem", "compile_surfaces", "affine_record"]


def affine_record(t: None | AbstractTransformation) -> _lib.Affine:
    a = _lib.Affine()
    if t is None:
        a.kind = 0
        a.m[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        return a
    af = t.affine()
    a.m[:] = [float(v) for v in af.matrix.ravel()]
    a.t[:] = [float(v) for v in af.vector]
    a.kind = 2 if af.has_rotation else 1
    return a


class _Tables:
    """Growable float64 table store with de-duplicated chemicals."""

    def __init__(self):
        self.data: list[np.ndarray] = []
        self.size = 0
        self.chemicals: list[tuple[int, int]] = []
        self._chem_index: dict[str, int] = {}
        self.layers: list[tuple[int, int, float, float]] = []
        self.stacks: list[tuple[int, int, list[tuple[int, int, int, int]]]] = []

    def add(self, arr) -> int:
        arr = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
        off = self.size
        self.data.append(arr)
        self.size += arr.size
        return off

    def chemical(self, chem) -> int:
        key = chem.file_nk
        if key not in self._chem_index:
            w, n, k = chem.nk
            off = self.add(np.concatenate([w, n, k]))
            self._chem_index[key] = len(self.chemicals)
            self.chemicals.append((off, len(w)))
        return self._chem_index[key]

    def layer(self, layer, thickness=None) -> int:
        chem = layer._chemical
        ci = -1 if chem is None else self.chemical(chem)
        prof = layer.interface
        pk = 0 if prof is None else int(prof.kind)
        pw = 0.0 if prof is None else float(prof.width)
        th = layer.thickness if thickness is None else thickness
        th = 0.0 if th is None else float(th)
        self.layers.append((ci, pk, th, pw))
        return len(self.layers) - 1

    def stack(self, layers, substrate) -> int:
        blocks: list[tuple[int, int, int, int]] = []
        seq: list[int] = []

        def flush():
            if seq:
                blocks.append((0, seq[0], len(seq), 1))
                seq.clear()

        def flat_layers(obj) -> list:
            if isinstance(obj, _mat.Layer):
                return [obj]
            if isinstance(obj, _mat.LayerSequence):
                out = []
                for x in obj.layers:
                    out += flat_layers(x)
                return out
            if isinstance(obj, (list, tuple)):
                out = []
                for x in obj:
                    out += flat_layers(x)
                return out
            raise TypeError(f"unsupported layer inside a period: {type(obj).__name__}")

        def walk(obj):
            if obj is None:
                return
            if isinstance(obj, _mat.Layer):
                # layers of one run must be contiguous in the layer table
                idx = self.layer(obj)
                if seq and idx != seq[-1] + 1:
                    flush()
                seq.append(idx)
            elif isinstance(obj, _mat.PeriodicLayerSequence):
                flush()
                period = flat_layers(list(obj.layers))
                first = len(self.layers)
                for L in period:
                    self.layer(L)
                blocks.append((1, first, len(period), int(obj.num_periods)))
            elif isinstance(obj, (_mat.LayerSequence, list, tuple)):
                items = obj.layers if isinstance(obj, _mat.LayerSequence) else obj
                for x in items:
                    walk(x)
            else:
                raise TypeError(f"unsupported layer type {type(obj).__name__}")

        walk(layers)
        flush()
        if len(blocks) > _lib.MAX_BLOCKS:
            raise ValueError(f"at most {_lib.MAX_BLOCKS} layer blocks are supported")
        sub = -1
        if substrate is not None:
            sub = self.layer(substrate, thickness=0.0)
        self.stacks.append((len(blocks), sub, blocks))
        return len(self.stacks) - 1

    def array(self) -> np.ndarray:
        if not self.data:
            return np.zeros(1)
        return np.concatenate(self.data)


def _orders_of(rulings) -> tuple[int, ...]:
    o = rulings.diffraction_order
    if np.ndim(o) == 0:
        return (int(o),)
    return tuple(int(v) for v in np.asarray(o).ravel())


def _surface_parts(surface):
    """(sag, material, aperture, rulings, transformation) of any surface,
    including an imaging sensor."""
    return (surface.sag, surface.material, surface.aperture,
            getattr(surface, "rulings", None), surface.transformation)


def _fill_surface(rec: _lib.Surface, surface, tables: _Tables, num_paths: int) -> None:
    sag, material, aperture, rulings, transformation = _surface_parts(surface)
    rec.transform = affine_record(transformation)

    # sag
    rec.sag_kind = int(sag.kind)
    rec.sag_p[:] = [float(v) for v in sag._params()]
    rec.sag_transform = affine_record(sag.transformation)

    # material
    rec.material_kind = int(material.kind)
    rec.is_mirror = 1 if material.is_mirror else 0
    rec.material_table = -1
    rec.stack = -1
    if isinstance(material, _mat.Glass):
        rec.material_p[:] = [float(material.b1), float(material.b2), float(material.b3),
                             float(material.c1) * 1e6, float(material.c2) * 1e6,
                             float(material.c3) * 1e6]
    elif isinstance(material, _mat.MeasuredMirror):
        w, e = material.efficiency_measured
        w = np.asarray(w, dtype=np.float64).ravel()
        e = np.asarray(e, dtype=np.float64).ravel()
        if w.size != e.size or w.size < 2:
            raise ValueError("efficiency_measured needs matching 1-D wavelength/efficiency arrays")
        rec.material_table = tables.add(np.concatenate([w, e]))
        rec.material_table_len = w.size
    elif isinstance(material, _mat.AbstractMultilayerMaterial):
        rec.stack = tables.stack(material.layers, getattr(material, "substrate", None))
    elif material.kind == _mat.MATERIAL_SILICON_SENSOR:
        # the oxide over the silicon substrate, and the substrate thickness
        # of the Beer-Lambert decay
        oxide, substrate = material._layers()
        rec.stack = tables.stack([oxide], substrate)
        rec.material_p[0] = float(material.thickness_substrate)

    # rulings
    rec.rulings_kind = 0
    rec.rulings_table = -1
    if rulings is not None:
        rec.rulings_kind = int(rulings.kind)
        orders = _orders_of(rulings)
        rec.num_orders = len(orders)
        for i, o in enumerate(orders):
            rec.orders[i] = o
        rec.rulings_p[:] = [float(v) for v in rulings._params()]
        sp = rulings.spacing_
        rec.spacing_kind = int(sp.kind)
        rec.wavelength_scale = float(sp.wavelength_scale)
        rec.spacing_transform = affine_record(getattr(sp, "transformation", None)
                                              if sp.kind == _rl.SPACING_POLY1D else None)
        if sp.kind in (_rl.SPACING_CONSTANT, _rl.SPACING_POLY1D):
            rec.spacing_normal[:] = [float(sp.normal.x), float(sp.normal.y), float(sp.normal.z)]
        if sp.kind == _rl.SPACING_CONSTANT:
            rec.spacing_constant = float(sp.constant)
        elif sp.kind == _rl.SPACING_POLY1D:
            rec.num_terms = len(sp.coefficients)
            for i, (power, coef) in enumerate(sp.coefficients.items()):
                rec.poly_power[i] = int(power)
                rec.poly_coeff[i] = float(coef)
        else:
            rec.holo_x1[:] = [float(sp.x1.x), float(sp.x1.y), float(sp.x1.z)]
            rec.holo_x2[:] = [float(sp.x2.x), float(sp.x2.y), float(sp.x2.z)]
            rec.holo_wavelength = float(sp.wavelength)
            rec.holo_diverging1 = 1 if sp.is_diverging_1 else 0
            rec.holo_diverging2 = 1 if sp.is_diverging_2 else 0
        if isinstance(rulings, _rl.MeasuredRulings):
            w, e = rulings.efficiency_measured
            w = np.asarray(w, dtype=np.float64).ravel()
            e = np.asarray(e, dtype=np.float64).ravel()
            rec.rulings_table = tables.add(np.concatenate([w, e]))
            rec.rulings_table_len = w.size

    # aperture
    rec.aperture_kind = 0
    rec.polygon_offset = -1
    if aperture is not None:
        rec.aperture_kind = int(aperture.kind)
        flags = 0
        if aperture.active:
            flags |= 1
        if aperture.inverted:
            flags |= 2
        if aperture.clips_direction:
            flags |= 4
        rec.aperture_flags = flags
        rec.aperture_p[:] = [float(v) for v in aperture._params()]
        rec.aperture_transform = affine_record(aperture.transformation)
        if aperture.kind == _ap.APERTURE_POLYGON:
            xy = aperture._vertices_xy()
            rec.polygon_offset = tables.add(np.concatenate([xy[:, 0], xy[:, 1]]))
            rec.polygon_count = len(xy)


def _num_paths(surfaces) -> int:
    count = 1
    for s in surfaces:
        rl = getattr(s, "rulings", None)
        if rl is None:
            continue
        k = len(_orders_of(rl))
        if k > _lib.MAX_ORDERS:
            raise ValueError(f"at most {_lib.MAX_ORDERS} diffraction orders are supported")
        if k > 1:
            if count not in (1, k):
                raise ValueError("surfaces carry diffraction-order axes of different lengths")
            count = k
    return count


@dataclasses.dataclass
class HostRecords:
    surfaces: ctypes.Array
    layers: ctypes.Array
    stacks: ctypes.Array
    chemicals: ctypes.Array
    tables: np.ndarray
    num_paths: int
    pre: list
    pre_is_inverse: list
    post: _lib.Affine

    def key(self) -> bytes:
        return b"".join([
            bytes(self.surfaces), bytes(self.layers), bytes(self.stacks), bytes(self.chemicals),
            self.tables.tobytes(), bytes(self.pre[0]), bytes(self.pre[1]),
            bytes((self.pre_is_inverse[0], self.pre_is_inverse[1])), bytes(self.post),
        ])


def build_records(surfaces: Sequence, pre=(None, None), pre_is_inverse=(0, 0),
                  post=None) -> HostRecords:
    surfaces = list(surfaces)
    tables = _Tables()
    num_paths = _num_paths(surfaces)
    recs = (_lib.Surface * max(len(surfaces), 1))()
    for i, s in enumerate(surfaces):
        _fill_surface(recs[i], s, tables, num_paths)
    layers = (_lib.Layer * max(len(tables.layers), 1))()
    for i, (ci, pk, th, pw) in enumerate(tables.layers):
        layers[i].chemical, layers[i].profile_kind = ci, pk
        layers[i].thickness, layers[i].profile_width = th, pw
    stacks = (_lib.Stack * max(len(tables.stacks), 1))()
    for i, (nb, sub, blocks) in enumerate(tables.stacks):
        stacks[i].num_blocks, stacks[i].substrate_layer = nb, sub
        for j, (kind, first, count, periods) in enumerate(blocks):
            b = stacks[i].blocks[j]
            b.kind, b.first_layer, b.num_layers, b.num_periods = kind, first, count, periods
    chems = (_lib.Chemical * max(len(tables.chemicals), 1))()
    for i, (off, cnt) in enumerate(tables.chemicals):
        chems[i].offset, chems[i].count = off, cnt
    rec = HostRecords(
        surfaces=recs, layers=layers, stacks=stacks, chemicals=chems, tables=tables.array(),
        num_paths=num_paths,
        pre=[affine_record(pre[0]), affine_record(pre[1])],
        pre_is_inverse=[int(pre_is_inverse[0]), int(pre_is_inverse[1])],
        post=affine_record(post),
    )
    rec.num_surfaces = len(surfaces)
    rec.num_layers = len(tables.layers)
    rec.num_stacks = len(tables.stacks)
    rec.num_chemicals = len(tables.chemicals)
    return rec


def system_desc(rec: HostRecords):
    """The ``optk_system_desc`` of host records, and the array it points at
    (keep it alive while the descriptor is in use)."""
    d = _lib.SystemDesc()
    d.surfaces = ctypes.cast(rec.surfaces, ctypes.POINTER(_lib.Surface))
    d.num_surfaces = rec.num_surfaces
    d.num_paths = rec.num_paths
    d.layers = ctypes.cast(rec.layers, ctypes.POINTER(_lib.Layer))
    d.num_layers = rec.num_layers
    d.stacks = ctypes.cast(rec.stacks, ctypes.POINTER(_lib.Stack))
    d.num_stacks = rec.num_stacks
    d.chemicals = ctypes.cast(rec.chemicals, ctypes.POINTER(_lib.Chemical))
    d.num_chemicals = rec.num_chemicals
    tab = np.ascontiguousarray(rec.tables, dtype=np.float64)
    d.tables = tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.num_tables = tab.size
    d.pre[0], d.pre[1] = rec.pre[0], rec.pre[1]
    d.pre_is_inverse[0], d.pre_is_inverse[1] = rec.pre_is_inverse
    d.post = rec.post
    return d, tab


class CompiledSystem:
    """A system uploaded to the device (owns an ``optk_system*``)."""

    def __init__(self, rec: HostRecords):
        L = _lib.lib()
        d, tab = system_desc(rec)
        handle = ctypes.c_void_p()
        _lib.check(L.optk_system_create(ctypes.byref(d), ctypes.byref(handle)))
        self.handle = handle
        self.records = rec
        self.num_paths = rec.num_paths
        self.num_surfaces = rec.num_surfaces
        self._key = rec.key()
        # the layout index of the specialised kernels it runs, or -1
        self.signature = int(L.optk_system_specialised(handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().optk_system_destroy(h)
            except Exception:
                pass
            self.handle = None


_CACHE_ATTR = "_optk_compiled"


def compile_surfaces(surfaces: Sequence, pre=(None, None), pre_is_inverse=(0, 0), post=None,
                     owner=None) -> CompiledSystem:
    """Compile (or reuse the cached compilation held by ``owner``)."""
    rec = build_records(surfaces, pre, pre_is_inverse, post)
    key = rec.key()
    if owner is not None:
        cache = getattr(owner, _CACHE_ATTR, None)
        if isinstance(cache, dict) and key in cache:
            return cache[key]
    cs = CompiledSystem(rec)
    if owner is not None:
        try:
            cache = getattr(owner, _CACHE_ATTR, None)
            if not isinstance(cache, dict):
                cache = {}
            if len(cache) > 8:
                cache.clear()
            cache[key] = cs
            object.__setattr__(owner, _CACHE_ATTR, cache)
        except (AttributeError, TypeError):
            pass
    return cs


def invariant_fields(surfaces: Sequence) -> bool:
    """True when no surface can change wavelength, index or attenuation of
    rays entering with index 1 and attenuation 0 (no glass)."""
    for s in surfaces:
        if isinstance(s.material, _mat.Glass):
            return False
    return True

