"""Detector exposure and the forward model on the device:

* optk_expose draws the same numbers as the host build of its per-pixel
  code (tests/harness, whose Philox and Poisson law are pinned in
  test_harness_random.py), up to last-bit differences of the device's
  transcendental functions at an accept/reject boundary; without noise it
  is the expectation exactly;
* ``SequentialSystem.image`` of a uniform scene through the stop-test
  Newtonian collects radiance x wavelength band x field solid angle x
  entrance pupil area (pi (40/3)^2 mm^2, the 10 mm stop imaged by the
  primary); the field's solid angle in steradians is the closed form
  dx (sin y1 - sin y0) of the angular square on the unit sphere."""

import numpy as np
import pytest
import torch

from harness import expose as host_expose
from optika_amd import _collect, workloads
from optika_amd.vectors import Cartesian2dVectorArray

pytestmark = pytest.mark.gpu


def test_expose_matches_host_draws():
    print("case: expose vs host", flush=True)
    rng = np.random.default_rng(3)
    mean = np.concatenate([rng.uniform(0, 10, 50_000), rng.uniform(10, 1e4, 50_000), [0.0, 1e7]])
    img = torch.as_tensor(mean, device="cuda")
    got = _collect.expose(img, 1.0, 0.0, noise=True, seed=99).cpu().numpy()
    want = host_expose(mean, 99)
    assert (got == want).mean() > 0.999
    assert np.all(got == np.floor(got)) and got[-2] == 0.0
    exact = _collect.expose(img, 2.5, 7.0, noise=False).cpu().numpy()
    np.testing.assert_array_equal(exact, 2.5 * mean)
    g = _collect.expose(torch.zeros(400_000, device="cuda", dtype=torch.float64), 1.0, 2.0,
                        noise=True, seed=5).cpu().numpy()
    assert abs(g.mean()) < 5 * 2.0 / np.sqrt(g.size) and abs(g.std() - 2.0) < 0.02


def test_image_of_a_uniform_scene():
    print("case: image forward model", flush=True)
    system = workloads.stop_newtonian()
    system.sensor.timedelta_exposure = 2.0
    wv = np.array([500.0, 520.0, 540.0])
    f = np.linspace(-0.03, 0.03, 7)
    R = np.full((2, 6, 6), 1e6)  # photons / (s nm sr mm^2)
    v = np.linspace(-1, 1, 31)
    img = system.image(wv, f, f, R, pupil=Cartesian2dVectorArray(v, v), noise=False, seed=4)
    omega = np.deg2rad(0.06) * 2 * np.sin(np.deg2rad(0.03))  # sr
    want = 2.0 * 1e6 * 40.0 * omega * np.pi * (10.0 / 0.75) ** 2
    total = float(img.sum())
    assert abs(total - want) < 0.02 * want, (total, want)
    per_band = system.image(wv, f, f, R, pupil=Cartesian2dVectorArray(v, v), noise=False, seed=4,
                            integrate=False)
    assert per_band.shape[0] == 2 and abs(float(per_band.sum()) - total) < 1e-9 * total
    noisy = system.image(wv, f, f, R, pupil=Cartesian2dVectorArray(v, v), noise=True, seed=4)
    assert np.all(noisy.cpu().numpy() == np.floor(noisy.cpu().numpy()))
    assert abs(float(noisy.sum()) - total) < 6 * np.sqrt(total)
