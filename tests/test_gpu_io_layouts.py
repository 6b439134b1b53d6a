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
