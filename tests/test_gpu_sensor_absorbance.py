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
