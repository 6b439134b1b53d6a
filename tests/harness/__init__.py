"""CPU build of the kernels' per-ray code, for tests only.

``optk_host.cpp`` includes the same headers the gfx950 kernels are built
from and is compiled here with g++ (-ffp-contract=off), so the device
arithmetic can be compared with the numpy oracle without a GPU.  The
library is built into ``tests/harness/_build`` on first use.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIB = HERE / "_build" / "optk_host.so"

_lib = None


def build() -> pathlib.Path:
    srcs = [HERE / "optk_host.cpp", ROOT / "optika_amd/csrc/optk_surface.h",
            ROOT / "optika_amd/csrc/optk_fast.h", ROOT / "optika_amd/csrc/optk_random.h",
            ROOT / "optika_amd/csrc/optk_physics.h", ROOT / "include/optk.h"]
    if LIB.exists() and LIB.stat().st_mtime > max(s.stat().st_mtime for s in srcs):
        return LIB
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(f".{os.getpid()}.tmp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    f"-I{ROOT / 'include'}", str(HERE / "optk_host.cpp"), "-o", str(tmp)],
                   check=True)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(str(build()))
    return _lib


def _desc(rec):
    from optika_amd import _lib as L

    d = L.SystemDesc()
    d.surfaces = ctypes.cast(rec.surfaces, ctypes.POINTER(L.Surface))
    d.num_surfaces = rec.num_surfaces
    d.num_paths = rec.num_paths
    d.layers = ctypes.cast(rec.layers, ctypes.POINTER(L.Layer))
    d.num_layers = rec.num_layers
    d.stacks = ctypes.cast(rec.stacks, ctypes.POINTER(L.Stack))
    d.num_stacks = rec.num_stacks
    d.chemicals = ctypes.cast(rec.chemicals, ctypes.POINTER(L.Chemical))
    d.num_chemicals = rec.num_chemicals
    tab = np.ascontiguousarray(rec.tables, dtype=np.float64)
    d.tables = tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.num_tables = tab.size
    d.pre[0], d.pre[1] = rec.pre[0], rec.pre[1]
    d.pre_is_inverse[0], d.pre_is_inverse[1] = rec.pre_is_inverse
    d.post = rec.post
    return d, tab


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def trace_grid(system, grid, sensor_frame: bool = True):
    """Host trace of a SequentialSystem: list over paths of dicts like the
    oracle's rays."""
    from optika_amd import _compile

    obj_t = system.object.transformation if system.object is not None else None
    post = system.sensor.transformation if sensor_frame else None
    rec = _compile.build_records(system.surfaces_all, pre=(obj_t, system.transformation),
                                 pre_is_inverse=(0, 1), post=post)
    d, tab = _desc(rec)
    n = grid.num
    P = rec.num_paths
    ins = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, float), (n,)))
           for v in (grid.field_x, grid.field_y, grid.pupil_x, grid.pupil_y, grid.wavelength)]
    inten = np.ones(n)
    outs = [np.empty(P * n) for _ in range(10)]
    mask = np.empty(P * n, dtype=np.uint8)
    arr = (ctypes.POINTER(ctypes.c_double) * 10)(*[_dp(o) for o in outs])
    L = lib()
    L.host_trace_grid.restype = ctypes.c_int
    L.host_trace_grid(ctypes.byref(d), *[_dp(a) for a in ins], _dp(inten), ctypes.c_int64(n),
                      ctypes.c_int(1 if system.object_is_at_infinity else 0),
                      ctypes.c_int(1 if post is not None else 0), arr,
                      mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    res = []
    for p in range(P):
        s = slice(p * n, (p + 1) * n)
        res.append(dict(
            position=np.stack([outs[0][s], outs[1][s], outs[2][s]]),
            direction=np.stack([outs[3][s], outs[4][s], outs[5][s]]),
            wavelength=outs[6][s], intensity=outs[7][s], attenuation=outs[8][s],
            index_refraction=outs[9][s], unvignetted=mask[s].astype(bool)))
    return res


def small_angle_sincos(x):
    """The float64 kernels' series sin and cos for |x| <= 0.01 rad."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    lib().host_small_angle_sincos(_dp(x), _dp(s), _dp(c), ctypes.c_int64(x.size))
    return s, c


def fast_skip_first(system, sensor_frame: bool = True) -> bool:
    """Whether the float32 form's grid kernels elide the system's surface 0
    (a pass-through plane the grid rays start on)."""
    from optika_amd import _compile

    obj_t = system.object.transformation if system.object is not None else None
    post = system.sensor.transformation if sensor_frame else None
    rec = _compile.build_records(system.surfaces_all, pre=(obj_t, system.transformation),
                                 pre_is_inverse=(0, 1), post=post)
    d, _ = _desc(rec)
    L = lib()
    L.host_fast_skip_first.restype = ctypes.c_int
    return bool(L.host_fast_skip_first(ctypes.byref(d)))


def trace_grid_f32(system, grid, sensor_frame: bool = True):
    """Host run of the float32 form (float32 inputs: the grid is rounded to
    float32 first).  Same result layout as :func:`trace_grid`."""
    from optika_amd import _compile

    obj_t = system.object.transformation if system.object is not None else None
    post = system.sensor.transformation if sensor_frame else None
    rec = _compile.build_records(system.surfaces_all, pre=(obj_t, system.transformation),
                                 pre_is_inverse=(0, 1), post=post)
    d, tab = _desc(rec)
    n = grid.num
    P = rec.num_paths
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    ins = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (n,)))
           for v in (grid.field_x, grid.field_y, grid.pupil_x, grid.pupil_y, grid.wavelength)]
    inten = np.ones(n, np.float32)
    outs = [np.empty(P * n, np.float32) for _ in range(10)]
    mask = np.empty(P * n, dtype=np.uint8)
    arr = (ctypes.POINTER(ctypes.c_float) * 10)(*[fp(o) for o in outs])
    L = lib()
    L.host_trace_grid_f32.restype = ctypes.c_int
    L.host_trace_grid_f32(ctypes.byref(d), *[fp(a) for a in ins], fp(inten), ctypes.c_int64(n),
                          ctypes.c_int(1 if system.object_is_at_infinity else 0),
                          ctypes.c_int(1 if post is not None else 0), arr,
                          mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    res = []
    for p in range(P):
        s = slice(p * n, (p + 1) * n)
        res.append(dict(
            position=np.stack([outs[0][s], outs[1][s], outs[2][s]]),
            direction=np.stack([outs[3][s], outs[4][s], outs[5][s]]),
            wavelength=outs[6][s], intensity=outs[7][s], attenuation=outs[8][s],
            index_refraction=outs[9][s], unvignetted=mask[s].astype(bool)))
    return res


def trace_rays_f32(surfaces, rays: dict):
    """Host run of the float32 form over a surface sequence; `rays` is an
    oracle-style ray dict (rounded to float32 on the way in)."""
    from optika_amd import _compile

    rec = _compile.build_records(list(surfaces))
    d, tab = _desc(rec)
    n = np.asarray(rays["wavelength"]).size
    P = rec.num_paths
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    b = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (n,)))  # noqa: E731
    ins = [b(c) for c in rays["position"]] + [b(c) for c in rays["direction"]] + [
        b(rays[k]) for k in ("wavelength", "intensity", "attenuation", "index_refraction")]
    outs = [np.empty(P * n, np.float32) for _ in range(10)]
    mask = np.empty(P * n, dtype=np.uint8)
    inp = (ctypes.POINTER(ctypes.c_float) * 10)(*[fp(a) for a in ins])
    arr = (ctypes.POINTER(ctypes.c_float) * 10)(*[fp(o) for o in outs])
    L = lib()
    L.host_trace_rays_f32.restype = ctypes.c_int
    L.host_trace_rays_f32(ctypes.byref(d), inp, ctypes.c_int64(n), arr,
                          mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    res = []
    for p in range(P):
        s = slice(p * n, (p + 1) * n)
        res.append(dict(
            position=np.stack([outs[0][s], outs[1][s], outs[2][s]]),
            direction=np.stack([outs[3][s], outs[4][s], outs[5][s]]),
            wavelength=outs[6][s], intensity=outs[7][s], attenuation=outs[8][s],
            index_refraction=outs[9][s], unvignetted=mask[s].astype(bool)))
    return res


def multilayer(layers, substrate, wavelength, cos, n):
    """Host multilayer (Rs, Rp, Ts, Tp) through the kernel code."""
    from optika_amd import _compile, materials, surfaces

    s = surfaces.Surface(material=materials.MultilayerMirror(layers=layers, substrate=substrate))
    rec = _compile.build_records([s])
    d, tab = _desc(rec)
    wl = np.ascontiguousarray(np.asarray(wavelength, float).ravel())
    m = wl.size
    c = np.broadcast_to(np.asarray(cos, complex), (m,))
    nn = np.broadcast_to(np.asarray(n, complex), (m,))
    cre, cim = np.ascontiguousarray(c.real), np.ascontiguousarray(c.imag)
    nre, nim = np.ascontiguousarray(nn.real), np.ascontiguousarray(nn.imag)
    outs = [np.empty(m) for _ in range(4)]
    lib().host_multilayer(ctypes.byref(d), ctypes.c_int(0), _dp(wl), _dp(cre), _dp(cim), _dp(nre),
                          _dp(nim), ctypes.c_int64(m), *[_dp(o) for o in outs])
    return outs


def grid_bracket(w, x):
    """Brackets of ``x`` in the ascending table ``w`` through the index grid
    of optk_physics.h: (j, steps walked, cell count)."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    j = np.empty(x.size, np.int32)
    steps = np.empty(x.size, np.int32)
    f = lib().host_grid_bracket
    f.restype = ctypes.c_int
    cells = f(_dp(w), ctypes.c_int(w.size), _dp(x), ctypes.c_int(x.size),
              j.ctypes.data_as(ctypes.c_void_p), steps.ctypes.data_as(ctypes.c_void_p))
    return j, steps, cells


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().host_philox(c, k, o)
    return list(o)


def expose(mean, seed, read_noise=0.0):
    m = np.ascontiguousarray(mean, dtype=np.float64)
    out = np.empty_like(m)
    f = lib().host_expose
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p]
    f(m.ctypes.data, m.size, seed, read_noise, out.ctypes.data)
    return out
