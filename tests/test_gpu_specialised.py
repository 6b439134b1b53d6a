"""Kernels specialised on a system's surface kinds (the Newtonian and the
spectrograph signatures) against the generic kernels, which read the kinds
from the records.  Float64: bit for bit (the library is compiled without
contraction, so fixing the kinds at compile time changes no operation).
Float32 form: this form lets the compiler fuse multiply-adds, and the
specialised code fuses differently, so outputs agree to float32 rounding
(two units in the last place of positions and directions, 1e-6 of
intensities) and masks on all but rays within rounding of an edge.  Grid
entry (sensor-frame rayfunction) and ray entry (propagate_rays with every
output field), with perturbed parameters (a rescaled instance keeps its
layout's kernel)."""

import numpy as np
import pytest
import torch

from optika_amd import _trace, workloads
from optika_amd.rays import RayVectorArray
from optika_amd.vectors import Cartesian3dVectorArray

pytestmark = pytest.mark.gpu


def _grid_inputs(g, dtype):
    f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device="cuda").to(dtype).contiguous()  # noqa: E731
    return [f(g.field_x), f(g.field_y), f(g.pupil_x), f(g.pupil_y), f(g.wavelength)]


def _equal(a, b, label):
    """Float64 rays bit for bit; float32 rays to float32 rounding."""
    exact = a.position.x.dtype == torch.float64
    fields = [("position", k) for k in range(3)] + [("direction", k) for k in range(3)]
    for name, k in fields:
        x, y = list(getattr(a, name))[k], list(getattr(b, name))[k]
        assert torch.equal(torch.isnan(x), torch.isnan(y)), f"{label} {name}[{k}] NaNs"
        m = ~torch.isnan(x)
        if exact:
            assert torch.equal(x[m], y[m]), f"{label} {name}[{k}]"
        else:
            x, y = x[m].double(), y[m].double()
            tol = (2e-6 if name == "position" else 2e-7) + 2.4e-7 * torch.abs(y)
            assert torch.all(torch.abs(x - y) <= tol), f"{label} {name}[{k}]"
    ma, mb = a.unvignetted, b.unvignetted
    if exact:
        assert torch.equal(ma, mb), f"{label} unvignetted"
    else:
        assert (ma == mb).double().mean().item() >= 0.99999, f"{label} unvignetted"
    x, y = a.intensity, b.intensity
    m = ~torch.isnan(x)
    if exact:
        assert torch.equal(x[m], y[m]), f"{label} intensity"
    else:
        rel = torch.abs(x[m].double() - y[m].double()) / torch.clamp(torch.abs(y[m].double()), min=1e-30)
        assert torch.all(rel <= 1e-6), f"{label} intensity"


CASES = {
    "newtonian": (workloads.newtonian, lambda n: workloads.newtonian_grid_random(n, seed=41)),
    "spectrograph": (workloads.spectrograph, lambda n: workloads.spectrograph_grid(n, seed=42)),
    "spectrograph_rescaled": (
        lambda: workloads.spectrograph(workloads.Perturbation(primary_radius=0.002,
                                                              grating_radii=(0.001, -0.001),
                                                              spacing=(0.004, -0.003, 0.002))),
        lambda n: workloads.spectrograph_grid(n, seed=43)),
    # primary and grating decentred and tilted: the primary's transformation
    # is a full affine map, the tilted-spectrograph signature
    "spectrograph_tilted": (
        lambda: workloads.spectrograph(workloads.Perturbation(primary_decenter=(0.08, -0.05),
                                                              primary_tilt=(0.02, -0.015),
                                                              grating_decenter=(-0.04, 0.06),
                                                              grating_tilt=(0.01, 0.012),
                                                              primary_radius=-0.001)),
        lambda n: workloads.spectrograph_grid(n, seed=45)),
}


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_specialised_grid_kernel_equals_generic(name, precision):
    make, grid = CASES[name]
    system = make()
    n = 300_001  # not a multiple of the block
    dtype = torch.float64 if precision == "f64" else torch.float32
    ins = _grid_inputs(grid(n), dtype)
    print(f"case: specialised grid kernel {name} {precision}", flush=True)
    spec = _trace.GridTracePlan(system, n, precision=precision)
    gen = _trace.GridTracePlan(system, n, precision=precision, generic=True)
    assert spec.cs.signature >= 0, "the system should match a specialised signature"
    a, b = spec(*ins), gen(*ins)
    torch.cuda.synchronize()
    if a.position.x.ndim == 2:
        for p in range(a.position.x.shape[0]):
            _equal(a[p], b[p], f"{name} path {p}")
    else:
        _equal(a, b, name)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_specialised_ray_kernel_equals_generic(precision):
    system = workloads.newtonian()
    n = 200_003
    g = workloads.newtonian_grid_random(n, seed=44)
    dtype = torch.float64 if precision == "f64" else torch.float32
    print(f"case: specialised ray kernel newtonian {precision}", flush=True)
    t = lambda a: torch.as_tensor(np.asarray(a), device="cuda", dtype=dtype)  # noqa: E731
    gen = torch.Generator(device="cuda").manual_seed(5)
    d = torch.randn(3, n, device="cuda", dtype=torch.float64, generator=gen) * 1e-3
    d[2] = 1.0
    d = (d / d.norm(dim=0)).to(dtype)
    zero = torch.zeros(n, device="cuda", dtype=dtype)
    rays = RayVectorArray(wavelength=t(g.wavelength),
                          position=Cartesian3dVectorArray(t(g.pupil_x), t(g.pupil_y), zero),
                          direction=Cartesian3dVectorArray(d[0], d[1], d[2]),
                          intensity=torch.ones_like(zero), attenuation=zero.clone(),
                          index_refraction=torch.ones_like(zero),
                          unvignetted=torch.ones(n, device="cuda", dtype=torch.bool))
    surfaces = system.surfaces_all
    a = _trace.trace_sequence(surfaces, rays, precision=precision)
    b = _trace.trace_sequence(surfaces, rays, precision=precision, generic=True)
    torch.cuda.synchronize()
    _equal(a, b, "rays")
