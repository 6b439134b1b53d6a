"""Attenuating rays through the device trace.

Two steps use a ray's attenuation alpha (1/mm):

* every sag step multiplies the intensity by exp(-alpha |displacement|)
  (``AbstractSag.propagate_rays``, optika/sags/_abc.py:109-122);
* a multilayer material (and the silicon sensor material) folds alpha into
  the ambient index, n + i alpha lambda / (4 pi) (optika/materials/
  _multilayers.py:924-926), which changes its reflectivity.

Float64: the HIP trace equals the oracle to 1e-12 (positions, directions,
intensities) with identical masks; a flat vacuum plane gives the closed
form I exp(-alpha L) for the path length L.  Float32: intensities within
1e-4 relative of the float64 trace, geometry as in the float32 zoo test.
"""

import numpy as np
import pytest
import torch

import oracle
from optika_amd import materials, propagators, sensors, surfaces
from optika_amd.rays import RayVectorArray
from optika_amd.transformations import Cartesian3dTranslation
from optika_amd.vectors import Cartesian3dVectorArray

from test_gpu_parity import ZOO, _zoo_rays, assert_parity, rays_to_dict, split_paths

pytestmark = pytest.mark.gpu


def _device_rays(r, dtype=torch.float64):
    f = lambda a: torch.as_tensor(np.asarray(a, np.float64), device="cuda").to(dtype)  # noqa: E731
    return RayVectorArray(
        wavelength=f(r["wavelength"]),
        position=Cartesian3dVectorArray(*(f(c) for c in r["position"])),
        direction=Cartesian3dVectorArray(*(f(c) for c in r["direction"])),
        intensity=f(r["intensity"]), attenuation=f(r["attenuation"]),
        index_refraction=f(r["index_refraction"]))


def _attenuating(r, lo, hi, seed):
    r = dict(r)
    r["attenuation"] = np.random.default_rng(seed).uniform(lo, hi, r["wavelength"].size)
    return r


ZOO_ATTENUATING = ["sphere_glass_circle", "conic_mirror_ellipse", "toroid_rotated_sag_octagon",
                   "rough_multilayer_mirror", "measured_grating_multilayer_film",
                   "sawtooth_poly_grating"]


@pytest.mark.parametrize("name", ZOO_ATTENUATING)
def test_zoo_with_attenuation_float64(name):
    """alpha in [0.001, 0.03] / mm: the path factors exp(-alpha L) of both
    hops and the ambient-index fold of the multilayers, against the oracle."""
    surf = ZOO[name]
    wl = (12.0, 70.0) if "multilayer" in name else (400.0, 700.0)
    r = _attenuating(_zoo_rays(wl=wl), 1e-3, 3e-2, 41)
    n = r["wavelength"].size
    print(f"case: attenuating zoo {name}", flush=True)
    after = surfaces.Surface(transformation=Cartesian3dTranslation(z=120.0))
    out = propagators.propagate_rays([surf, after], _device_rays(r))
    ref = oracle.propagate([surf, after], r)
    gpu = split_paths(out, n, len(ref))
    for p in range(len(ref)):
        assert_parity(gpu[p], ref[p], f"{name} path {p}")
        live = ref[p]["unvignetted"] & np.isfinite(ref[p]["intensity"])
        assert np.all(ref[p]["intensity"][live] < r["intensity"][live])  # attenuated


def _near_surface_rays(z_surface, n, wl, alpha, seed):
    """Rays that start 1e-4 mm in front of a surface at z_surface, so that a
    strong attenuation (alpha L ~ 1) leaves them finite intensities."""
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(-5, 5, n), rng.uniform(-5, 5, n), np.full(n, z_surface - 1e-4)])
    d = oracle.direction_from_angles(rng.uniform(-8, 8, n), rng.uniform(-8, 8, n))
    return oracle.make_rays(rng.uniform(*wl, n), pos, d, intensity=rng.uniform(0.5, 1, n),
                            attenuation=rng.uniform(*alpha, n), index_refraction=1.0)


def _silicon_surface(z):
    return surfaces.Surface(
        material=sensors.BackIlluminatedSiliconSensorMaterial(
            thickness_oxide=5.0, thickness_implant=231.7, thickness_substrate=7000.0,
            roughness_oxide=0.5, roughness_substrate=1.0, cce_backsurface=0.21),
        transformation=Cartesian3dTranslation(z=z))


STRONG = {
    "multilayer_mirror": (lambda: surfaces.Surface(
        material=materials.MultilayerMirror(
            layers=materials.PeriodicLayerSequence(
                [materials.Layer("Si", thickness=4.02), materials.Layer("Mo", thickness=2.68)], 20),
            substrate=materials.Layer("SiO2")),
        transformation=Cartesian3dTranslation(z=10.0)), (12.0, 20.0)),
    "silicon_sensor": (lambda: _silicon_surface(10.0), (400.0, 700.0)),
}


@pytest.mark.parametrize("name", sorted(STRONG))
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_strong_attenuation_changes_the_coating(name, precision):
    """alpha in [1e3, 2e4] / mm: the ambient index gains an imaginary part
    alpha lambda / (4 pi) of 1e-3 .. 3e-2 (float64 against the oracle at
    1e-12; float32 within 1e-4 of the float64 trace), and the efficiency
    differs from that of the same rays without attenuation."""
    make, wl = STRONG[name]
    surf = make()
    r = _near_surface_rays(10.0, 4000, wl, (1e3, 2e4), 43)
    n = r["wavelength"].size
    print(f"case: strong attenuation {name} {precision}", flush=True)
    dt = torch.float64 if precision == "f64" else torch.float32
    if precision == "f32":
        # the float32 form reads float32 inputs: both forms see those values
        r = {k: (np.asarray(v, np.float32).astype(np.float64) if k != "unvignetted" else v)
             for k, v in r.items()}
    out = propagators.propagate_rays([surf], _device_rays(r, dt), precision=precision)
    ref = oracle.propagate([surf], r)[0]
    got = rays_to_dict(out, n)
    if precision == "f64":
        assert_parity(got, ref, name)
    else:
        assert np.array_equal(got["unvignetted"], ref["unvignetted"])
        rel = np.abs(got["intensity"] - ref["intensity"]) / ref["intensity"]
        assert rel.max() < 1e-4, rel.max()
    r0 = dict(r)
    r0["attenuation"] = np.zeros(n)
    ref0 = oracle.propagate([surf], r0)[0]
    path = np.exp(-r["attenuation"] * np.sqrt(((ref["position"] - r["position"]) ** 2).sum(0)))
    change = ref["intensity"] / (ref0["intensity"] * path) - 1
    assert np.median(np.abs(change)) > 1e-3  # the coating saw the absorbing ambient


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_flat_plane_closed_form(precision):
    """A vacuum plane at z = 25 mm reached from z = 0: intensity
    I exp(-alpha 25 / d_z) (float64: 1e-13 relative; float32: 1e-6)."""
    n = 2048
    rng = np.random.default_rng(47)
    d = oracle.direction_from_angles(rng.uniform(-30, 30, n), rng.uniform(-30, 30, n))
    pos = np.stack([rng.uniform(-5, 5, n), rng.uniform(-5, 5, n), np.zeros(n)])
    alpha = rng.uniform(0, 0.2, n)
    r = oracle.make_rays(np.full(n, 500.0), pos, d, intensity=np.full(n, 3.0), attenuation=alpha)
    dt = torch.float64 if precision == "f64" else torch.float32
    if precision == "f32":
        r = {k: (np.asarray(v, np.float32).astype(np.float64) if k != "unvignetted" else v)
             for k, v in r.items()}
    plane = surfaces.Surface(transformation=Cartesian3dTranslation(z=25.0))
    print(f"case: attenuation closed form {precision}", flush=True)
    out = propagators.propagate_rays([plane], _device_rays(r, dt), precision=precision)
    got = out.intensity.double().cpu().numpy()
    want = 3.0 * np.exp(-r["attenuation"] * 25.0 / r["direction"][2])
    np.testing.assert_allclose(got, want, rtol=1e-13 if precision == "f64" else 1e-6)
    assert got.min() < 0.01 * 3.0 and got.max() > 0.5 * 3.0
