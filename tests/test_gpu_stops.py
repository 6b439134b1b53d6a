"""The stop solve on the device (optika_amd/stops.py) against the oracle's
(oracle/stops.py, pinned by the reference's known answers and closed forms
in tests/test_oracle_stops.py), and ``raytrace`` / ``rayfunction`` with
normalized field and pupil coordinates (the reference's default) against the
oracle trace of the oracle-denormalized grid."""

import numpy as np
import pytest
import torch

from oracle import stops as S
from oracle import trace as T
from optika_amd.vectors import Cartesian2dVectorArray
from stop_systems import RADIUS_FIELD_GRAZING, RADIUS_FIELD_NEWTONIAN, grazing, newtonian

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make,radius", [(newtonian, RADIUS_FIELD_NEWTONIAN),
                                         (grazing, RADIUS_FIELD_GRAZING)])
def test_field_and_pupil_extent_match_oracle(make, radius):
    print(f"case: stop extent {make.__name__}", flush=True)
    system = make()
    (flo, fhi), (plo, phi) = S.field_pupil_extent(system, [500.0])
    got = [system.field_min, system.field_max, system.pupil_min, system.pupil_max]
    want = [flo, fhi, plo, phi]
    # both solves stop once every residual is below 1e-9 of the target
    # outline's extent; the device stops each ray on its own, the oracle
    # when all have converged, so the roots agree to that tolerance
    for g, w in zip(got, want):
        np.testing.assert_allclose(np.stack([g.x, g.y]), w, rtol=0, atol=2e-7)
    assert np.all(np.abs(np.asarray(system.field_max.x) - radius) < 1e-6)


@pytest.mark.parametrize("make", [newtonian, grazing])
def test_normalized_rayfunction_matches_oracle(make):
    print(f"case: normalized rayfunction {make.__name__}", flush=True)
    system = make()
    rng = np.random.default_rng(5)
    n = 4000
    u = rng.uniform(-1, 1, (4, n))
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
    out = system.rayfunction(wavelength=500.0, field=Cartesian2dVectorArray(t(u[0]), t(u[1])),
                             pupil=Cartesian2dVectorArray(t(u[2]), t(u[3])))
    fx, fy, px, py = S.denormalize(system, 500.0, u[0], u[1], u[2], u[3])
    ref = T.rayfunction(system, 500.0, fx, fy, px, py)[0]
    r = out.outputs
    np.testing.assert_allclose(out.inputs.field.x.cpu().numpy(), fx, rtol=0, atol=1e-8)
    np.testing.assert_allclose(out.inputs.pupil.y.cpu().numpy(), py, rtol=0, atol=2e-7)
    m = r.unvignetted.cpu().numpy()
    agree = (m == ref["unvignetted"]).mean()
    assert agree > 0.999  # rays within ~1e-9 mm of an edge may differ
    both = m & ref["unvignetted"]
    for k, c in enumerate((r.position.x, r.position.y, r.position.z)):
        np.testing.assert_allclose(c.cpu().numpy()[both], ref["position"][k][both], rtol=0, atol=2e-6)


def test_default_grid_and_per_ray_wavelengths():
    """No arguments: ``grid_input`` read as normalized (the reference's
    ``_rayfunction_input``); per-ray wavelengths map each ray through its
    own wavelength's extent."""
    from optika_amd.vectors import ObjectVectorArray

    print("case: default grid / per-ray wavelengths", flush=True)
    system = grazing()
    dev = torch.device("cuda")
    lin = torch.linspace(-1, 1, 5, device=dev, dtype=torch.float64)
    gx, gy = torch.meshgrid(lin, lin, indexing="ij")
    system.grid_input = ObjectVectorArray(
        wavelength=500.0, field=Cartesian2dVectorArray(gx.reshape(-1), gy.reshape(-1)),
        pupil=Cartesian2dVectorArray(gx.reshape(-1) * 0.5, gy.reshape(-1) * 0.5))
    r = system.rayfunction()
    fx = r.inputs.field.x.cpu().numpy()
    assert abs(fx.max() - RADIUS_FIELD_GRAZING) < 1e-6 and abs(fx.min() + RADIUS_FIELD_GRAZING) < 1e-6
    wl = torch.tensor([500.0, 600.0] * 10, device=dev, dtype=torch.float64)
    f = Cartesian2dVectorArray(torch.ones(20, device=dev, dtype=torch.float64),
                               torch.zeros(20, device=dev, dtype=torch.float64))
    p = Cartesian2dVectorArray(torch.ones(20, device=dev, dtype=torch.float64),
                               torch.zeros(20, device=dev, dtype=torch.float64))
    r2 = system.raytrace(wavelength=wl, field=f, pupil=p, accumulate=False)
    (_, fhi), (_, phi) = S.field_pupil_extent(system, [500.0, 600.0])
    got = r2.inputs.pupil.x.cpu().numpy()
    np.testing.assert_allclose(got[0::2], phi[0, 0], atol=2e-7)
    np.testing.assert_allclose(got[1::2], phi[0, 1], atol=2e-7)


def test_missing_stop_raises():
    from optika_amd import stops

    system = newtonian()
    system.surfaces[1].is_pupil_stop = False
    with pytest.raises(ValueError, match="upil"):
        stops.rayfunction_stops(system, [500.0])


def test_device_newton_roots_land_on_the_outline():
    """The solved rays, traced forward by the oracle, meet the pupil stop's
    outline points within the solve's tolerance (1e-9 of 20 mm)."""
    from optika_amd import stops

    print("case: device Newton roots", flush=True)
    system = newtonian()
    r = stops.solve_stop_rays(system, [500.0])
    surf = [s for s, _, _ in S.surfaces_all(system)]
    rays = T.make_rays(500.0, r["position"].reshape(3, -1), r["direction"].reshape(3, -1))
    out = T.propagate(surf[1:3], rays)[0]
    local = T.transform_inverse(surf[2].transformation, out["position"])[0]
    want = S.wire(surf[2].aperture, 21)[:2]
    got = local[:2].reshape(2, 21, 21)
    assert np.all(np.abs(got - want[:, None, :]) < 2e-8 + 1e-12)


def test_spectrograph_normalized_rayfunction_per_order():
    """The spectrograph's orders -1 and +1 both lie between its pupil stop
    (the primary) and its field stop: the reference's stop solve carries the
    order axis, so each order has its own field and pupil extent and the
    normalized grid is denormalized per order (_sequential.py:396-789).
    Each path's denormalized inputs and traced rays match the oracle's
    solve and trace of that path."""
    from optika_amd import stops, workloads

    print("case: spectrograph normalized rayfunction, both orders", flush=True)
    system = workloads.spectrograph()
    assert stops.num_stop_paths(system) == 2
    rng = np.random.default_rng(6)
    n = 3000
    u = rng.uniform(-1, 1, (4, n))
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
    out = system.rayfunction(wavelength=60.0, field=Cartesian2dVectorArray(t(u[0]), t(u[1])),
                             pupil=Cartesian2dVectorArray(t(u[2]), t(u[3])))
    r = out.outputs
    assert r.position.x.shape == (2, n)
    assert out.inputs.field.x.shape == (2, n)
    ext = stops.field_pupil(system, stops.rayfunction_stops(system, [60.0]))
    for p in range(2):
        (flo, fhi), (plo, phi) = S.field_pupil_extent(system, [60.0], path=p)
        for v, lo, hi in ((ext[0], flo, fhi), (ext[1], plo, phi)):
            np.testing.assert_allclose(v[:, p].min(axis=(1, 2)), lo, rtol=0, atol=2e-7)
            np.testing.assert_allclose(v[:, p].max(axis=(1, 2)), hi, rtol=0, atol=2e-7)
        fx, fy, px, py = S.denormalize(system, 60.0, u[0], u[1], u[2], u[3], path=p)
        np.testing.assert_allclose(out.inputs.field.x[p].cpu().numpy(), fx, rtol=0, atol=1e-8)
        np.testing.assert_allclose(out.inputs.field.y[p].cpu().numpy(), fy, rtol=0, atol=1e-8)
        np.testing.assert_allclose(out.inputs.pupil.x[p].cpu().numpy(), px, rtol=0, atol=2e-7)
        np.testing.assert_allclose(out.inputs.pupil.y[p].cpu().numpy(), py, rtol=0, atol=2e-7)
        ref = T.rayfunction(system, 60.0, fx, fy, px, py)[p]
        m = r.unvignetted[p].cpu().numpy()
        assert (m == ref["unvignetted"]).mean() > 0.999  # rays within ~1e-9 of an edge may differ
        # order -1 reaches the sensor; order +1, whose field the stops allow,
        # is clipped downstream of the field stop in the reference as here
        assert ref["unvignetted"].mean() > (0.3 if p == 0 else -1.0)
        for k, c in enumerate((r.position.x[p], r.position.y[p], r.position.z[p])):
            g, w = c.cpu().numpy(), ref["position"][k]
            fin = np.isfinite(g) & np.isfinite(w)
            assert fin.mean() > 0.99
            np.testing.assert_allclose(g[fin], w[fin], rtol=0, atol=2e-6)
    # the extents differ between the orders (the +1 order sees another field)
    assert abs(ext[0][0, 0].min() - ext[0][0, 1].min()) > 1.0


def test_device_roots_on_a_tilted_decentred_stop_match_the_host_newton():
    """A perturbed spectrograph (primary tilted and decentred, the first
    stop of the solve): the device's per-ray Newton stopping rule (each ray
    stops when its own residual is below tolerance) and the oracle's
    whole-array rule (every ray iterates until all are below it) give roots
    within 1e-9 of each other in direction cosine, per order path."""
    from optika_amd import stops, workloads

    print("case: tilted/decentred stop roots vs host Newton", flush=True)
    pert = workloads.Perturbation(primary_tilt=(0.01, -0.012), primary_decenter=(0.05, -0.03),
                                  primary_radius=0.001, grating_tilt=(0.0, 0.0),
                                  grating_decenter=(0.0, 0.0), grating_radii=(0.0, 0.0),
                                  spacing=(0.0, 0.0, 0.0))
    system = workloads.spectrograph(pert)
    surfaces = system.surfaces_all
    for p in range(2):
        got = stops.solve_stop_rays(system, [59.0, 62.0], surfaces=stops.single_path(surfaces, p))
        want = S.stop_rays(system, [59.0, 62.0], path=p)
        np.testing.assert_allclose(got["position"], want["position"], rtol=1e-12, atol=1e-12)
        err = np.abs(got["direction"] - want["direction"]).max()
        print(f"path {p}: max direction difference {err:.2e}", flush=True)
        assert err < 1e-9
