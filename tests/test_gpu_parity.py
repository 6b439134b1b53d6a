"""Parity of the HIP float64 trace (through the C ABI) with the CPU oracle.

Tolerance (float64): every position and direction component within
1e-12 * max(|oracle|, 1) -- 1e-12 relative, with a 1 mm (or 1, for
direction cosines) floor for components near zero -- and identical
vignetting masks.  Intensities (multilayer transfer matrices) within
1e-12 relative.  Sizes are small enough for the oracle to run in seconds;
the full 16M-ray workload is checked on a seeded random subsample and by
size-independent properties.
"""

import numpy as np
import pytest
import torch

import oracle
from optika_amd import (
    apertures,
    materials,
    propagators,
    rulings,
    sags,
    sensors,
    surfaces,
    workloads,
)
from optika_amd.rays import RayVectorArray
from optika_amd.transformations import (
    Cartesian3dRotationX,
    Cartesian3dRotationY,
    Cartesian3dRotationZ,
    Cartesian3dTranslation,
    TransformationList,
)
from optika_amd.vectors import Cartesian2dVectorArray, Cartesian3dVectorArray

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _np(t):
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return np.asarray(t)


def rays_to_dict(r: RayVectorArray, n: int) -> dict:
    def f(v, dt=np.float64):
        return np.broadcast_to(np.asarray(_np(v), dtype=dt), (n,)).copy()

    return dict(
        position=np.stack([f(c) for c in r.position]),
        direction=np.stack([f(c) for c in r.direction]),
        wavelength=f(r.wavelength), intensity=f(r.intensity), attenuation=f(r.attenuation),
        index_refraction=f(r.index_refraction), unvignetted=f(r.unvignetted, bool))


def assert_parity(gpu: dict, ref: dict, label=""):
    assert np.array_equal(gpu["unvignetted"], ref["unvignetted"]), f"{label}: masks differ"
    for k in ("position", "direction"):
        a, b = gpu[k], ref[k]
        both_nan = np.isnan(a) & np.isnan(b)
        same_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
        ok = both_nan | same_inf | (np.abs(a - b) <= RTOL * np.maximum(np.abs(b), 1.0))
        assert ok.all(), f"{label} {k}: max err {np.nanmax(np.abs(a - b)[~ok])}"
    for k in ("wavelength", "intensity", "attenuation", "index_refraction"):
        a, b = gpu[k], ref[k]
        ok = (np.isnan(a) & np.isnan(b)) | (a == b) | (np.abs(a - b) <= RTOL * np.abs(b))
        assert ok.all(), f"{label} {k}: max rel err {np.nanmax(np.abs(a - b)[~ok] / np.abs(b[~ok]))}"


def gpu_rayfunction(system, g, **kw):
    out = system.rayfunction(wavelength=torch.as_tensor(g.wavelength, device="cuda"),
                             field=Cartesian2dVectorArray(torch.as_tensor(g.field_x, device="cuda"),
                                                          torch.as_tensor(g.field_y, device="cuda")),
                             pupil=Cartesian2dVectorArray(torch.as_tensor(g.pupil_x, device="cuda"),
                                                          torch.as_tensor(g.pupil_y, device="cuda")),
                             normalized_field=False, normalized_pupil=False, **kw)
    return out.outputs


def split_paths(r: RayVectorArray, n: int, paths: int) -> list[dict]:
    if paths == 1:
        return [rays_to_dict(r, n)]
    return [rays_to_dict(r[p], n) for p in range(paths)]


def test_newtonian_c1_full_grid():
    g = workloads.newtonian_grid_c1()
    s = workloads.newtonian()
    print("case: C1 newtonian, 20480 rays", flush=True)
    gpu = split_paths(gpu_rayfunction(s, g), g.num, 1)
    ref = oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y)
    assert_parity(gpu[0], ref[0], "C1")


def test_newtonian_16m_subsample_and_properties():
    """The full 16M-ray Newtonian workload: a seeded subsample of 20k rays
    matches the oracle; every ray satisfies size-independent invariants."""
    n = 1 << 24
    g = workloads.newtonian_grid_random(n, seed=1)
    s = workloads.newtonian()
    print("case: C2 newtonian, 16M rays", flush=True)
    out = gpu_rayfunction(s, g)
    idx = np.random.default_rng(7).choice(n, 20000, replace=False)
    sub = g.subset(0)
    sub.field_x, sub.field_y = g.field_x[idx], g.field_y[idx]
    sub.pupil_x, sub.pupil_y, sub.wavelength = g.pupil_x[idx], g.pupil_y[idx], g.wavelength[idx]
    ref = oracle.rayfunction(s, sub.wavelength, sub.field_x, sub.field_y, sub.pupil_x, sub.pupil_y)[0]
    it = torch.as_tensor(idx, device="cuda")
    gpu = rays_to_dict(RayVectorArray(
        wavelength=out.wavelength[it] if isinstance(out.wavelength, torch.Tensor) else out.wavelength,
        position=Cartesian3dVectorArray(*(c[it] for c in out.position)),
        direction=Cartesian3dVectorArray(*(c[it] for c in out.direction)),
        intensity=out.intensity[it], attenuation=out.attenuation,
        index_refraction=out.index_refraction, unvignetted=out.unvignetted[it]), len(idx))
    assert_parity(gpu, ref, "C2 subsample")
    # invariants over all 16M rays
    d = torch.stack(list(out.direction))
    norm = torch.sqrt((d * d).sum(0))
    assert torch.max(torch.abs(norm - 1)).item() < 1e-14
    u = out.unvignetted
    x, y = out.position.x[u], out.position.y[u]
    assert torch.all(torch.abs(x) <= 15.36) and torch.all(torch.abs(y) <= 15.36)
    # the vignetted fraction over all 16M rays agrees with the oracle's
    # subsample within 5 standard errors
    frac = u.double().mean().item()
    f_ref = ref["unvignetted"].mean()
    assert abs(frac - f_ref) < 5 * np.sqrt(f_ref * (1 - f_ref) / len(idx)) + 1e-9


def test_spectrograph_both_orders():
    g = workloads.spectrograph_grid(20000, seed=2)
    s = workloads.spectrograph()
    print("case: C3 spectrograph, 20000 rays x 2 orders", flush=True)
    gpu = split_paths(gpu_rayfunction(s, g), g.num, 2)
    ref = oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y)
    for p in range(2):
        assert_parity(gpu[p], ref[p], f"C3 path {p}")


def test_raytrace_accumulate_global_frame():
    g = workloads.newtonian_grid_random(4000, seed=5)
    s = workloads.newtonian()
    print("case: raytrace accumulate", flush=True)
    out = s.raytrace(wavelength=torch.as_tensor(g.wavelength, device="cuda"),
                     field=Cartesian2dVectorArray(g.field_x, g.field_y),
                     pupil=Cartesian2dVectorArray(g.pupil_x, g.pupil_y),
                     normalized_field=False, normalized_pupil=False, accumulate=True).outputs
    ref = oracle.raytrace(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y,
                          accumulate=True)[0]
    assert out.position.x.shape == (4, g.num)
    for k in range(4):
        assert_parity(rays_to_dict(out[k], g.num), ref[k], f"surface {k}")


# -- the operator zoo: every sag, ruling, material and aperture ---------------

def _surface_zoo():
    tilt = TransformationList([Cartesian3dRotationX(3.0), Cartesian3dRotationY(-2.0),
                               Cartesian3dTranslation(x=1.0, y=-2.0, z=50.0)])
    cases = {
        "sphere_glass_circle": surfaces.Surface(
            sag=sags.SphericalSag(radius=-150.0), material=materials.Glass.n_bk7(),
            aperture=apertures.CircularAperture(20.0), transformation=Cartesian3dTranslation(z=40.0)),
        "conic_mirror_ellipse": surfaces.Surface(
            sag=sags.ConicSag(radius=-400.0, conic=-0.6, transformation=Cartesian3dTranslation(x=3.0)),
            material=materials.Mirror(), aperture=apertures.EllipticalAperture(radius=(15.0, 10.0)),
            transformation=tilt),
        "parabola_measured_mirror_sector": surfaces.Surface(
            sag=sags.ParabolicSag(focal_length=-300.0),
            material=materials.MeasuredMirror(efficiency_measured=(np.linspace(400, 700, 7),
                                                                   np.linspace(0.5, 0.9, 7))),
            aperture=apertures.CircularSectorAperture(radius=25.0, angle_start=-120.0, angle_stop=150.0),
            transformation=Cartesian3dTranslation(z=30.0)),
        "cylinder_vacuum_polygon": surfaces.Surface(
            sag=sags.CylindricalSag(radius=300.0),
            aperture=apertures.PolygonalAperture(vertices=[[-10, -8, 0], [12, -6, 0], [8, 9, 0],
                                                           [-3, 4, 0], [-11, 10, 0]]),
            transformation=Cartesian3dTranslation(z=25.0)),
        "toroid_rotated_sag_octagon": surfaces.Surface(
            sag=sags.ToroidalSag(radius=200.0, radius_of_rotation=350.0,
                                 transformation=Cartesian3dRotationZ(20.0)),
            material=materials.Mirror(), aperture=apertures.OctagonalAperture(radius=18.0),
            transformation=Cartesian3dTranslation(z=60.0)),
        "flat_trapezoid_inverted": surfaces.Surface(
            aperture=apertures.IsoscelesTrapezoidalAperture(x_left=-5.0, x_right=12.0, angle=30.0,
                                                            inverted=True),
            transformation=Cartesian3dTranslation(z=10.0)),
        "sinusoidal_grating": surfaces.Surface(
            material=materials.Mirror(),
            rulings=rulings.SinusoidalRulings(spacing=2e-3, depth=30.0, diffraction_order=[-1, 0, 1]),
            transformation=Cartesian3dTranslation(z=20.0)),
        "square_grating": surfaces.Surface(
            material=materials.Mirror(),
            rulings=rulings.SquareRulings(spacing=2e-3, depth=30.0, diffraction_order=[0, 1, 3]),
            transformation=Cartesian3dTranslation(z=20.0)),
        "sawtooth_poly_grating": surfaces.Surface(
            sag=sags.SphericalSag(radius=-500.0), material=materials.Mirror(),
            rulings=rulings.SawtoothRulings(
                spacing=rulings.Polynomial1dRulingSpacing(coefficients={0: 1e-3, 1: 1e-6, 2: 1e-8, 3: 1e-10},
                                                          normal=(1.0, 0.0, 0.0),
                                                          transformation=Cartesian3dRotationZ(10.0)),
                depth=40.0, diffraction_order=[1, 2]),
            transformation=Cartesian3dTranslation(z=20.0)),
        "triangular_grating": surfaces.Surface(
            material=materials.Mirror(),
            rulings=rulings.TriangularRulings(spacing=1.5e-3, depth=25.0, diffraction_order=[-2, -1]),
            transformation=Cartesian3dTranslation(z=20.0)),
        "rectangular_holographic_grating": surfaces.Surface(
            material=materials.Mirror(),
            rulings=rulings.RectangularRulings(
                spacing=rulings.HolographicRulingSpacing(x1=(5.0, 0.0, -10.0), x2=(10.0, 0.0, -10.0),
                                                         wavelength=500.0),
                depth=20.0, ratio_duty=0.3, diffraction_order=[0, 1]),
            transformation=Cartesian3dTranslation(z=20.0)),
        "measured_grating_multilayer_film": surfaces.Surface(
            material=materials.MultilayerFilm(layers=[materials.Layer("SiO2", thickness=5.0),
                                                      materials.Layer("Cr", thickness=2.0)]),
            rulings=rulings.MeasuredRulings(spacing=3e-3, diffraction_order=1,
                                            efficiency_measured=(np.linspace(300, 800, 11),
                                                                 np.linspace(0.1, 0.3, 11))),
            transformation=Cartesian3dTranslation(z=15.0)),
        "rough_multilayer_mirror": surfaces.Surface(
            sag=sags.SphericalSag(radius=-800.0),
            material=materials.MultilayerMirror(
                layers=[materials.Layer("SiC", thickness=25.0,
                                        interface=materials.profiles.ErfInterfaceProfile(1.0)),
                        materials.PeriodicLayerSequence(
                            [materials.Layer("Si", thickness=4.0,
                                             interface=materials.profiles.LinearInterfaceProfile(0.5)),
                             materials.Layer("Mo", thickness=2.7,
                                             interface=materials.profiles.SinusoidalInterfaceProfile(0.4))],
                            num_periods=10),
                        materials.Layer("Cr", thickness=3.0,
                                        interface=materials.profiles.ExponentialInterfaceProfile(0.6))],
                substrate=materials.Layer("SiO2")),
            transformation=Cartesian3dTranslation(z=35.0)),
    }
    return cases


ZOO = _surface_zoo()


def _zoo_rays(n=3000, seed=9, wl=(400.0, 700.0)):
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), np.zeros(n)])
    ang = rng.uniform(-5, 5, (2, n))
    d = oracle.direction_from_angles(ang[0], ang[1])
    return oracle.make_rays(rng.uniform(*wl, n), pos, d, intensity=rng.uniform(0.5, 1, n),
                            attenuation=0.0, index_refraction=1.0)


@pytest.mark.parametrize("name", sorted(ZOO))
def test_operator_zoo_propagate(name):
    surf = ZOO[name]
    wl = (12.0, 70.0) if "multilayer" in name else (400.0, 700.0)
    r = _zoo_rays(wl=wl)
    n = r["wavelength"].size
    print(f"case: zoo {name}", flush=True)
    rays = RayVectorArray(
        wavelength=torch.as_tensor(r["wavelength"], device="cuda"),
        position=Cartesian3dVectorArray(*(torch.as_tensor(c, device="cuda") for c in r["position"])),
        direction=Cartesian3dVectorArray(*(torch.as_tensor(c, device="cuda") for c in r["direction"])),
        intensity=torch.as_tensor(r["intensity"], device="cuda"))
    after = surfaces.Surface(transformation=Cartesian3dTranslation(z=120.0))
    out = propagators.propagate_rays([surf, after], rays)
    ref = oracle.propagate([surf, after], r)
    paths = len(ref)
    gpu = split_paths(out, n, paths)
    for p in range(paths):
        assert_parity(gpu[p], ref[p], f"{name} path {p}")


def test_sensor_collect_matches_histogramdd():
    g = workloads.newtonian_grid_random(200000, seed=21)
    s = workloads.newtonian()
    print("case: collect 200k rays", flush=True)
    out = gpu_rayfunction(s, g)
    edges = np.array([400.0, 500.0, 600.0, 700.0])
    image, cosine = s.sensor.collect(out, edges)
    ref_rays = oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y)[0]
    ref_image, ref_cos = oracle.collect(s.sensor, ref_rays, edges)
    img = image.cpu().numpy()
    assert img.shape == ref_image.shape == (3, 2048, 2048)
    assert np.array_equal(img > 0, ref_image > 0)
    np.testing.assert_allclose(img, ref_image, rtol=1e-12, atol=0)
    np.testing.assert_allclose(cosine.cpu().numpy(), ref_cos, rtol=1e-12, atol=0)


def test_sensor_collect_point_sources():
    """The C1 grid: five field points, each imaged onto a few pixels, so
    whole wavefronts of rays share a bin (the wave-combined atomics path)."""
    g = workloads.newtonian_grid_c1()
    s = workloads.newtonian()
    print("case: collect C1 point sources", flush=True)
    out = gpu_rayfunction(s, g)
    edges = np.array([400.0, 700.0])
    image, cosine = s.sensor.collect(out, edges)
    ref_rays = oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y)[0]
    ref_image, ref_cos = oracle.collect(s.sensor, ref_rays, edges)
    img = image.cpu().numpy()
    assert np.array_equal(img > 0, ref_image > 0)
    assert (img > 0).sum() < 200  # a few pixels per field point
    np.testing.assert_allclose(img, ref_image, rtol=1e-12, atol=0)
    np.testing.assert_allclose(cosine.cpu().numpy(), ref_cos, rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("n,edges", [(1 << 20, [400.0, 500.0, 600.0, 700.0]),
                                     (40_000, [400.0, 700.0]), (32_769, [400.0, 700.0]),
                                     (3, [400.0, 700.0])])
def test_collect_tiled_equals_atomic_path(n, edges, precision):
    """The tiled counting-sort binning and the global-atomic binning give the
    same image (rounding of the order of additions only), for float64 and
    float32 rays, also for ragged ray counts that leave the last block's
    chunk (32768 rays) nearly or partly empty."""
    from optika_amd import _collect

    g = workloads.newtonian_grid_random(n, seed=31)
    s = workloads.newtonian()
    print(f"case: collect tiled vs atomic n={n} {precision}", flush=True)
    out = gpu_rayfunction(s, g, precision=precision)
    hw = 15e-3 * 2048 / 2
    a, ma = _collect.histogram(out, edges, (-hw, hw), (-hw, hw), 2048, 2048, tiled=True)
    b, mb = _collect.histogram(out, edges, (-hw, hw), (-hw, hw), 2048, 2048, tiled=False)
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.array_equal(a > 0, b > 0)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=0)
    np.testing.assert_allclose(ma.cpu().numpy(), mb.cpu().numpy(), rtol=1e-12, atol=1e-300)
    assert abs(a.sum() - out.intensity[out.unvignetted].double().sum().item()) <= 1e-9 * max(a.sum(), 1)


# -- per-operator entry points --------------------------------------------------

def test_eval_operators_match_oracle():
    rng = np.random.default_rng(31)
    n = 2000
    p = np.stack([rng.uniform(-30, 30, n), rng.uniform(-30, 30, n), rng.uniform(-1, 1, n)])
    pt = Cartesian3dVectorArray(*(torch.as_tensor(c, device="cuda") for c in p))
    print("case: eval operators", flush=True)
    for sag in (sags.SphericalSag(radius=-200.0), sags.ConicSag(radius=300.0, conic=-0.3),
                sags.ParabolicSag(focal_length=-120.0), sags.CylindricalSag(radius=150.0),
                sags.ToroidalSag(radius=250.0, radius_of_rotation=400.0), sags.NoSag()):
        z = sag(pt).cpu().numpy()
        np.testing.assert_allclose(z, oracle.sag_value(sag, p), rtol=1e-13, atol=1e-13)
        nrm = sag.normal(pt)
        np.testing.assert_allclose(np.stack([_np(c) for c in nrm]), oracle.sag_normal(sag, p),
                                   rtol=1e-13, atol=1e-13)
    for ap in (apertures.CircularAperture(20.0), apertures.RectangularAperture(half_width=(15.0, 8.0)),
               apertures.RegularPolygonalAperture(radius=22.0, num_vertices=6),
               apertures.EllipticalAperture(radius=(25.0, 12.0))):
        m = ap(pt).cpu().numpy()
        assert np.array_equal(m, oracle.aperture_mask(ap, p)), type(ap).__name__
    # multilayer_efficiency as a free function
    layers = materials.PeriodicLayerSequence(
        [materials.Layer("Si", thickness=4.02), materials.Layer("Mo", thickness=2.68)], 40)
    wl = np.linspace(12.0, 20.0, 101)
    R, T = materials.multilayer_efficiency(torch.as_tensor(wl, device="cuda"), 0.98, 1.0, layers,
                                           materials.Layer("SiO2"))
    (oRs, oRp), (oTs, oTp) = oracle.multilayer_efficiency(wl, 0.98, 1.0, layers, materials.Layer("SiO2"))
    np.testing.assert_allclose(_np(R.s), oRs, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(_np(R.p), oRp, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(_np(T.s), oTs, rtol=1e-12, atol=1e-15)
    # vector Snell's law
    d = oracle.direction_from_angles(rng.uniform(-20, 20, n), rng.uniform(-20, 20, n))
    dt = Cartesian3dVectorArray(*(torch.as_tensor(c, device="cuda") for c in d))
    for mirror in (False, True):
        b = materials.snells_law(dt, 1.0, 1.5, is_mirror=mirror)
        ref = oracle.snells_law(d, 1.0, 1.5, np.array([[0.0], [0.0], [-1.0]]), mirror)
        np.testing.assert_allclose(np.stack([_np(c) for c in b]), ref, rtol=1e-13, atol=1e-15)


# -- edge cases -------------------------------------------------------------------

def test_empty_input():
    s = workloads.newtonian()
    print("case: empty input", flush=True)
    out = s.rayfunction(wavelength=torch.zeros(0, dtype=torch.float64, device="cuda"),
                        field=Cartesian2dVectorArray(torch.zeros(0, dtype=torch.float64, device="cuda"), 0.0),
                        pupil=Cartesian2dVectorArray(torch.zeros(0, dtype=torch.float64, device="cuda"), 0.0),
                        normalized_field=False, normalized_pupil=False).outputs
    assert out.position.x.numel() == 0


def test_rays_that_miss_and_total_internal_reflection():
    """No conic root -> t = inf (optika/sags/_conic.py:154); total internal
    reflection -> NaN direction (_snells_law.py:362).  Both propagate as in
    the reference, and the masks agree."""
    miss = surfaces.Surface(sag=sags.ConicSag(radius=-10.0, conic=0.0), material=materials.Mirror(),
                            aperture=apertures.CircularAperture(5.0),
                            transformation=Cartesian3dTranslation(z=5.0))
    exit_face = surfaces.Surface(material=materials.Vacuum(),
                                 aperture=apertures.CircularAperture(50.0),
                                 transformation=Cartesian3dTranslation(z=5.0))
    n = 64
    pos = np.stack([np.linspace(-30, 30, n), np.zeros(n), np.zeros(n)])
    d = oracle.direction_from_angles(np.linspace(-60, 60, n), np.zeros(n))
    print("case: miss + TIR", flush=True)
    rays = RayVectorArray(wavelength=500.0,
                          position=Cartesian3dVectorArray(*(torch.as_tensor(c, device="cuda") for c in pos)),
                          direction=Cartesian3dVectorArray(*(torch.as_tensor(c, device="cuda") for c in d)),
                          index_refraction=1.6)
    r = oracle.make_rays(500.0, pos, d, index_refraction=1.6)
    for seq in ([miss], [exit_face]):
        out = propagators.propagate_rays(seq, rays)
        ref = oracle.propagate(seq, r)[0]
        assert_parity(rays_to_dict(out, n), ref, "edge")
    assert np.isinf(oracle.propagate([miss], r)[0]["position"]).any()
    assert np.isnan(oracle.propagate([exit_face], r)[0]["direction"]).any()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_collect_narrow_offset_range_matches_histogramdd(precision):
    """Pixel ranges narrow for their distance from the origin (num * (|lo| +
    |hi|) / (hi - lo) far above 2^48, where a first-guess bin can be off by
    many bins and numpy's edges repeat) bin exactly as numpy.histogramdd with
    numpy.linspace edges does: every ray in the same pixel (unit weights, so
    the sums are exact counts), including rays on edges and on the last edge."""
    from optika_amd import _collect

    rng = np.random.default_rng(17)
    nx, ny = 2048, 64
    # ~1.4e17 by the ratio above: 2048 bins over ~260 distinct float64
    # values, so most edges repeat; one multiply and two correction steps
    # would put ~90 % of these rays in the wrong bin
    x_lo, x_hi = 1e6, 1e6 + 3e-8
    y_lo, y_hi = -3e7, -3e7 + 1.0  # ~4e9: the first guess would still do
    n = 200_000
    x_edges = np.linspace(x_lo, x_hi, nx + 1)
    y_edges = np.linspace(y_lo, y_hi, ny + 1)
    dt = np.float64 if precision == "f64" else np.float32
    # points inside the ranges, plus exact edge values (representable in the
    # rays' precision, so cast the edges first)
    x = rng.uniform(x_lo - 1.5e-9, x_hi + 1.5e-9, n)
    y = rng.uniform(y_lo - 0.01, y_hi + 0.01, n)
    x[:4096] = rng.choice(x_edges, 4096)
    y[:4096] = rng.choice(y_edges, 4096)
    x, y = x.astype(dt).astype(np.float64), y.astype(dt).astype(np.float64)
    wl = np.full(n, 550.0)
    edges = np.array([400.0, 700.0])
    t = lambda a: torch.as_tensor(a.astype(dt), device="cuda")  # noqa: E731
    print(f"case: collect narrow offset range {precision}", flush=True)
    rays = RayVectorArray(wavelength=t(wl), position=Cartesian3dVectorArray(t(x), t(y), t(np.zeros(n))),
                          direction=Cartesian3dVectorArray(t(np.zeros(n)), t(np.zeros(n)), t(np.ones(n))),
                          intensity=t(np.ones(n)))
    for tiled in (True, False):
        img, mom = _collect.histogram(rays, edges, (x_lo, x_hi), (y_lo, y_hi), nx, ny, tiled=tiled)
        want, _ = np.histogramdd(np.stack([wl, x, y], 1), bins=(edges, x_edges, y_edges))
        got = img.cpu().numpy()
        if precision == "f64":
            # most rays inside, over many pixels (float32 coordinates cannot
            # resolve such a range: they collapse onto a few values, binned
            # the same way all the same)
            assert want.sum() > 0.8 * n and (want > 0).sum() > 5000
        assert np.array_equal(got, want), f"tiled={tiled}: {np.abs(got - want).sum()} rays misbinned"
