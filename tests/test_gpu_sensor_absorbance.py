"""Silicon sensor absorbance (optika_amd.sensors.transmittance / absorbance,
the reference's optika/sensors/materials/_materials.py:55-376) on the
device multilayer kernel:

* the reference's own checks (_materials_test.py:41-117): transmittance and
  absorbance lie in [0, 1] for its wavelengths (304 A; 100-200 A), normal and
  oblique incidence;
* a closed form: without the oxide, at normal incidence from vacuum, the
  transmittance into the substrate is 4 Re(n) / |1 + n|^2 and the
  absorbance that times 1 - exp(-4 pi Im(n) d / lambda)."""

import numpy as np
import pytest
import torch

from optika_amd import sensors
from optika_amd.chemicals import Chemical

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wavelength", [np.array([30.4]), np.linspace(10.0, 20.0, 4)])
@pytest.mark.parametrize("direction", [1.0, np.cos(np.deg2rad(30.0))])
def test_ranges(wavelength, direction):
    print(f"case: absorbance range {wavelength} {direction}", flush=True)
    w = torch.as_tensor(wavelength, device="cuda")
    t = sensors.transmittance(w, direction=direction)
    a = sensors.absorbance(w, direction=direction)
    for v in (t.s, t.p):
        v = v.cpu().numpy()
        assert np.all(v >= 0) and np.all(v <= 1)
    for v in (a.s, a.p):
        assert np.all(v >= 0) and np.all(v <= 1)


def test_closed_form_without_oxide():
    print("case: absorbance closed form", flush=True)
    wl = np.array([13.5, 30.4, 121.6, 500.0])
    w = torch.as_tensor(wl, device="cuda")
    d = 2000.0
    t = sensors.transmittance(w, thickness_oxide=0.0, thickness_substrate=d)
    a = sensors.absorbance(w, thickness_oxide=0.0, thickness_substrate=d)
    n = Chemical("Si").n(wl)
    t_want = 4 * n.real / np.abs(1 + n) ** 2
    a_want = t_want * (1 - np.exp(-4 * np.pi * n.imag * d / wl))
    for v in (t.s, t.p):
        np.testing.assert_allclose(v.cpu().numpy(), t_want, rtol=1e-10)
    for v in (a.s, a.p):
        np.testing.assert_allclose(v, a_want, rtol=1e-10, atol=1e-300)


def _newtonian_with_silicon():
    from optika_amd import workloads

    s = workloads.newtonian()
    s.sensor.material = sensors.BackIlluminatedSiliconSensorMaterial(
        thickness_oxide=5.0, thickness_implant=231.7, thickness_substrate=7000.0,
        roughness_oxide=0.5, roughness_substrate=1.0, cce_backsurface=0.21)
    return s


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_silicon_sensor_in_the_trace(precision):
    """The sensor material's efficiency applied per ray inside the trace
    kernels: the float64 form equals the oracle (intensity 1e-12, geometry
    1e-12, identical masks); the float32 form's intensities are within 1e-4
    of the oracle's."""
    import oracle
    from optika_amd import workloads
    from test_gpu_parity import gpu_rayfunction

    g = workloads.newtonian_grid_random(50_000, seed=23)
    s = _newtonian_with_silicon()
    print(f"case: silicon sensor trace {precision}", flush=True)
    out = gpu_rayfunction(s, g, precision=precision)
    o = oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y)[0]
    m = out.unvignetted.cpu().numpy()
    assert np.array_equal(m, o["unvignetted"])
    got = out.intensity.double().cpu().numpy()
    if precision == "f64":
        np.testing.assert_allclose(got, o["intensity"], rtol=1e-12, atol=1e-15)
        px = out.position.x.cpu().numpy()
        np.testing.assert_allclose(px[m], o["position"][0][m], rtol=1e-12, atol=1e-12)
    else:
        rel = np.abs(got[m] - o["intensity"][m]) / o["intensity"][m]
        assert rel.max() < 1e-4


def test_silicon_sensor_efficiency_matches_its_absorbance():
    """At normal incidence the per-ray efficiency (the trace's material
    step, device eval) equals the material's own absorbance average (the
    host composition of the device multilayer and Snell kernels)."""
    from optika_amd.rays import RayVectorArray
    from optika_amd.vectors import Cartesian3dVectorArray

    m = _newtonian_with_silicon().sensor.material
    wl = np.linspace(400.0, 700.0, 7)
    print("case: silicon sensor efficiency vs absorbance", flush=True)
    w = torch.as_tensor(wl, device="cuda")
    one = torch.ones_like(w)
    zero = torch.zeros_like(w)
    rays = RayVectorArray(wavelength=w, position=Cartesian3dVectorArray(zero, zero, zero),
                          direction=Cartesian3dVectorArray(zero, zero, one), intensity=one,
                          attenuation=zero, index_refraction=one)
    eff = m.efficiency(rays, Cartesian3dVectorArray(0.0, 0.0, -1.0)).cpu().numpy()
    a = m.absorbance(w)
    want = (np.asarray(a.s) + np.asarray(a.p)) / 2
    np.testing.assert_allclose(eff, want, rtol=1e-12)
