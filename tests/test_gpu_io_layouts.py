"""The two I/O forms of the grid kernels give identical rays.

``optk_trace_grid`` launches a kernel specialised for the sensor-frame
rayfunction layout (field, pupil and wavelength arrays in, the intensity a
single value; position, direction, intensity and mask out) whenever the
descriptors match it, and the general kernel, which tests every pointer,
otherwise.  Passing the same intensity as an array instead of a value
selects the general kernel on the same rays; both forms must then agree bit
for bit, in both precisions, for the geometric (Newtonian) and the full
feature set (spectrograph: toroid, rulings, multilayers, two orders).
"""

import pytest
import torch

from optika_amd import _trace, workloads

pytestmark = pytest.mark.gpu

CASES = [
    ("newtonian", workloads.newtonian, workloads.newtonian_grid_random, 50_000),
    ("spectrograph", workloads.spectrograph, workloads.spectrograph_grid, 20_000),
]


def _fields(rays):
    out = {
        "px": rays.position.x, "py": rays.position.y, "pz": rays.position.z,
        "dx": rays.direction.x, "dy": rays.direction.y, "dz": rays.direction.z,
        "intensity": rays.intensity, "unvignetted": rays.unvignetted,
    }
    return {k: v.detach().reshape(-1).cpu() for k, v in out.items()}


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name,make_system,make_grid,n", CASES, ids=[c[0] for c in CASES])
def test_sensor_layout_equals_general_layout(name, make_system, make_grid, n, precision):
    system = make_system()
    grid = make_grid(n, seed=5).object_vector(device="cuda")
    print(f"io layouts: {name} {precision} n={n}", flush=True)
    sensor = _trace.trace_system(system, grid, intensity=None, sensor_frame=True,
                                 precision=precision).outputs
    ones = torch.ones(n, dtype=torch.float64, device="cuda")
    general = _trace.trace_system(system, grid, intensity=ones, sensor_frame=True,
                                  precision=precision).outputs
    torch.cuda.synchronize()
    a, b = _fields(sensor), _fields(general)
    assert a["unvignetted"].any()
    for k in a:
        assert a[k].shape == b[k].shape, k
        same = a[k] == b[k]
        if a[k].is_floating_point():
            same |= torch.isnan(a[k]) & torch.isnan(b[k])
        assert bool(same.all()), f"{name} {precision}: {k} differs between I/O forms"


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name,make_system,make_grid,n", CASES, ids=[c[0] for c in CASES])
def test_ray_entry_full_outputs_equal_general_layout(name, make_system, make_grid, n, precision):
    """The ray entry point (``optk_trace``) launches a kernel that writes
    every output field without null tests when all are requested (as
    ``propagate_rays`` does), and the general kernel otherwise.  Leaving the
    index output out selects the general kernel on the same rays: every
    other field must agree bit for bit."""
    import ctypes

    from optika_amd import _lib
    from optika_amd.rays import RayVectorArray
    from optika_amd.vectors import Cartesian3dVectorArray

    system = make_system()
    g = make_grid(n, seed=6)
    print(f"io layouts (rays): {name} {precision} n={n}", flush=True)
    dtype = torch.float64 if precision == "f64" else torch.float32
    # rays in the system frame on the object plane, near the axis
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    d = torch.randn(3, n, device="cuda", dtype=torch.float64, generator=gen) * 1e-3
    d[2] = 1.0
    d = (d / d.norm(dim=0)).to(dtype)
    t = lambda a: torch.as_tensor(a, device="cuda", dtype=dtype)  # noqa: E731
    zero = torch.zeros(n, device="cuda", dtype=dtype)
    first = RayVectorArray(wavelength=t(g.wavelength),
                           position=Cartesian3dVectorArray(t(g.pupil_x), t(g.pupil_y), zero),
                           direction=Cartesian3dVectorArray(d[0], d[1], d[2]),
                           intensity=torch.ones_like(zero), attenuation=zero.clone(),
                           index_refraction=torch.ones_like(zero),
                           unvignetted=torch.ones(n, device="cuda", dtype=torch.bool))
    full = _trace.trace_sequence(system.surfaces_all, first, precision=precision)
    cs = _trace.compiled(system.surfaces_all)
    P = cs.num_paths
    out = _trace._Out(P * n, torch.device("cuda"), dtype, {"wavelength", "intensity", "attenuation"})
    rin = _lib.Rays()
    keep = []
    for k, c in enumerate(first.position):
        c = c.contiguous()
        keep.append(c)
        rin.position[k] = c.data_ptr()
    for k, c in enumerate(first.direction):
        c = c.contiguous()
        keep.append(c)
        rin.direction[k] = c.data_ptr()
    for f in ("wavelength", "intensity", "attenuation", "index_refraction"):
        v = getattr(first, f).contiguous()
        keep.append(v)
        setattr(rin, f, v.data_ptr())
    u = first.unvignetted.to(torch.uint8).contiguous()
    keep.append(u)
    rin.unvignetted = u.data_ptr()
    prec = _lib.F64 if precision == "f64" else _lib.F32
    _lib.check(_lib.lib().optk_trace(cs.handle, prec, ctypes.byref(rin), ctypes.byref(out.rec), None,
                                     n, 0, _trace.stream_handle()))
    torch.cuda.synchronize()
    a = _fields(full)
    b = {"px": out.t["position"][0], "py": out.t["position"][1], "pz": out.t["position"][2],
         "dx": out.t["direction"][0], "dy": out.t["direction"][1], "dz": out.t["direction"][2],
         "intensity": out.t["intensity"], "unvignetted": out.t["unvignetted"]}
    b = {k: v.detach().reshape(-1).cpu() for k, v in b.items()}
    assert a["unvignetted"].any()
    for k in a:
        same = a[k] == b[k]
        if a[k].is_floating_point():
            same |= torch.isnan(a[k]) & torch.isnan(b[k])
        assert bool(same.all()), f"{name} {precision}: {k} differs between I/O forms"
