"""Pin the CPU oracle against the reference's own known answers, golden
files and invariants (optika/**/_tests), and against closed forms and brute
force where those leave a function open.  CPU only."""

import math
import pathlib

import numpy as np
import pytest
import scipy.optimize

import oracle
from optika_amd import apertures, materials, rulings, sags, sensors, surfaces, systems
from optika_amd.transformations import (
    Cartesian3dRotationY,
    Cartesian3dRotationZ,
    Cartesian3dTranslation,
    TransformationList,
)
from optika_amd.vectors import Cartesian2dVectorArray

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"

# the reference's transformation fixtures (optika/_tests/test_mixins.py:12-28)
TRANSFORMS = [
    None,
    Cartesian3dTranslation(x=5.0),
    TransformationList([Cartesian3dTranslation(x=5.0), Cartesian3dRotationZ(53.0),
                        Cartesian3dTranslation(x=6.0)]),
]


def ref_rays():
    """The deterministic ray fixtures of optika/rays/_tests/test_ray_vectors.py:11-38."""
    r0 = oracle.make_rays(500.0, [0.0, 1.0, 2.0], [0.0, 0.0, 1.0], attenuation=0.0)
    n = 11
    ang = np.deg2rad(np.linspace(-1, 1, n))
    lin = np.linspace(-10, 10, n)
    r1 = oracle.make_rays(np.linspace(400, 500, n), np.stack([lin, lin, lin]),
                          np.stack([np.sin(ang), np.zeros(n), np.cos(ang)]),
                          attenuation=np.linspace(0, 1, n),
                          index_refraction=np.linspace(1, 1.5, n))
    return [r0, r1]


def _sag_cases():
    cases = []
    for R in (100.0, -100.0, 500.0):
        for T in TRANSFORMS:
            cases.append(sags.SphericalSag(radius=R, transformation=T))
            for k in (0.0, 0.5, 1.0, -0.5):
                cases.append(sags.ConicSag(radius=R, conic=k, transformation=T))
            cases.append(sags.ParabolicSag(focal_length=R / 2, transformation=T))
            cases.append(sags.CylindricalSag(radius=R, transformation=T))
            for Rr in (200.0, -200.0):
                if R > 0 and Rr > 0:
                    cases.append(sags.ToroidalSag(radius=R, radius_of_rotation=Rr,
                                                  transformation=T))
        cases.append(sags.NoSag())
    return cases


SAG_CASES = _sag_cases()


def _id(s):
    t = type(s.transformation).__name__ if s.transformation is not None else "None"
    return f"{type(s).__name__}-{t}-{getattr(s, 'radius', '')}-{getattr(s, 'conic', '')}"


@pytest.mark.parametrize("sag", SAG_CASES, ids=[_id(s) for s in SAG_CASES])
@pytest.mark.parametrize("ray_index", [0, 1])
def test_intercept_on_surface_and_matches_root_finder(sag, ray_index):
    """optika/sags/_tests/_abc_test.py:90-102: the intercept lies on the sag and
    equals a generic root of line(t).z - sag(line(t)); the generic root is
    found independently here with scipy's Brent method."""
    rays = ref_rays()[ray_index]
    res = oracle.sag_intercept(sag, rays)
    pos = res["position"]
    # np.allclose defaults, as the reference's test (its parabola formula
    # cancels digits for near-axis rays)
    assert np.allclose(oracle.sag_value(sag, pos), pos[2])
    o, u = rays["position"], rays["direction"]
    for i in range(o.shape[1]):
        def f(t):
            p = (o[:, i] + u[:, i] * t)[:, None]
            return p[2, 0] - oracle.sag_value(sag, p)[0]

        t_closed = np.dot(pos[:, i] - o[:, i], u[:, i])
        # bracket the root near the closed-form distance
        lo, hi = t_closed - 1.0, t_closed + 1.0
        if np.sign(f(lo)) == np.sign(f(hi)):
            continue  # grazing tangency: no sign change to bracket
        t_root = scipy.optimize.brentq(f, lo, hi, xtol=1e-14, rtol=1e-15)
        assert abs(t_root - t_closed) <= 1e-8 + 1e-5 * abs(t_closed)


@pytest.mark.parametrize("sag", [s for s in SAG_CASES if type(s).__name__ != "NoSag"],
                         ids=[_id(s) for s in SAG_CASES if type(s).__name__ != "NoSag"])
def test_normal_unit_and_downward(sag):
    """optika/sags/_tests/_abc_test.py:78-88."""
    rng = np.random.default_rng(0)
    p = np.stack([rng.uniform(-5, 5, 50), rng.uniform(-5, 5, 50), np.zeros(50)])
    n = oracle.sag_normal(sag, p)
    assert np.all(n[2] < 0)
    assert np.allclose(np.sqrt(n[0] ** 2 + n[1] ** 2 + n[2] ** 2), 1)


def test_normal_matches_numerical_gradient():
    """The analytic normals equal the normalised gradient of z - sag(x, y)."""
    rng = np.random.default_rng(1)
    p = np.stack([rng.uniform(-20, 20, 40), rng.uniform(-20, 20, 40), np.zeros(40)])
    for sag in [sags.SphericalSag(radius=-300.0), sags.ConicSag(radius=250.0, conic=-0.7),
                sags.ParabolicSag(focal_length=-150.0), sags.CylindricalSag(radius=400.0),
                sags.ToroidalSag(radius=300.0, radius_of_rotation=500.0)]:
        h = 1e-5
        gx = (oracle.sag_value(sag, p + [[h], [0], [0]]) - oracle.sag_value(sag, p - [[h], [0], [0]])) / (2 * h)
        gy = (oracle.sag_value(sag, p + [[0], [h], [0]]) - oracle.sag_value(sag, p - [[0], [h], [0]])) / (2 * h)
        v = np.stack([gx, gy, -np.ones_like(gx)])
        v = v / np.sqrt((v ** 2).sum(0))
        assert np.allclose(oracle.sag_normal(sag, p), v, atol=1e-8), type(sag).__name__


def test_parabolic_normal_equals_conic_normal():
    """optika/sags/_tests/_parabolic_test.py:25-37."""
    p = np.stack([np.linspace(-50, 50, 11), np.linspace(-30, 30, 11), np.zeros(11)])
    a = sags.ParabolicSag(focal_length=-100.0)
    b = sags.ConicSag(radius=-200.0, conic=-1.0)
    assert np.allclose(oracle.sag_normal(a, p), oracle.sag_normal(b, p))


def test_intercept_grazing_conic():
    """optika/sags/_tests/_conic_test.py:57-87 (same inputs)."""
    sag = sags.ConicSag(radius=-0.7, conic=-1.0007)
    az = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    rays = oracle.make_rays(500.0, np.stack([68 * np.cos(az), 68 * np.sin(az), np.full(64, -2000.0)]),
                            [0.0, 0.0, 1.0])
    pos = oracle.sag_intercept(sag, rays)["position"]
    assert np.allclose(oracle.sag_value(sag, pos), pos[2])
    c = 1 / sag.radius
    r2 = pos[0] ** 2 + pos[1] ** 2
    assert np.all((pos[2] * (c * r2 - pos[2])) >= 0)


def test_toroid_with_equal_radii_is_a_sphere():
    """A torus whose two radii agree is a sphere: the Newton intercept must
    agree with the sphere's closed form."""
    rng = np.random.default_rng(2)
    n = 200
    o = np.stack([rng.uniform(-30, 30, n), rng.uniform(-30, 30, n), np.full(n, -50.0)])
    d = np.stack([rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n), np.ones(n)])
    d = d / np.sqrt((d ** 2).sum(0))
    rays = oracle.make_rays(500.0, o, d)
    R = 400.0
    a = oracle.sag_intercept(sags.ToroidalSag(radius=R, radius_of_rotation=R), rays)["position"]
    b = oracle.sag_intercept(sags.SphericalSag(radius=R), rays)["position"]
    assert np.allclose(a, b, rtol=0, atol=1e-11)


# -- Snell's law ----------------------------------------------------------------


@pytest.mark.parametrize("is_mirror", [False, True])
@pytest.mark.parametrize("n2", [1.5, 1.0, 1.333, 2.0])
@pytest.mark.parametrize("angles", [(0.0, 0.0), (1.0, 2.0)])
def test_snells_law_closed_form(is_mirror, n2, angles):
    """optika/materials/_tests/test_snells_law.py:82-120."""
    a = oracle.direction_from_angles(np.array([angles[0]]), np.array([angles[1]]))
    normal = np.array([[0.0], [0.0], [-1.0]])
    n1 = 1.0
    b = oracle.snells_law(a, n1, n2, normal, is_mirror)
    r = n1 / n2
    c = -(a * normal).sum(0)
    mirror = np.sign(c) * (2 * is_mirror - 1)
    t = np.sqrt(np.square(1 / r) + np.square(c) - (a * a).sum(0))
    expected = r * (a + (c + mirror * t) * normal)
    assert np.allclose(np.sqrt((b * b).sum(0)), 1)
    same_side = np.sign((a * normal).sum(0)) == np.sign((b * normal).sum(0))
    assert np.all(same_side != is_mirror)
    assert np.allclose(b, expected)


@pytest.mark.parametrize("n2", [1.5, 1.5 + 0.2j, 1.0, 2.0])
def test_snells_law_scalar_closed_form(n2):
    """optika/materials/_tests/test_snells_law.py:29-46."""
    cos = np.linspace(-1, 1, 7)
    res = oracle.snells_law_scalar(cos, 1, n2)
    expected = np.cos(np.emath.arcsin(1 * np.sin(np.emath.arccos(cos)) / n2))
    assert np.allclose(res, expected)


# -- materials ------------------------------------------------------------------


def test_glass_sellmeier_known_indices():
    """optika/materials/_tests/test_materials.py:126-150: n_d of N-BK7 and F2."""
    wl = {"d": 587.5618, "F": 486.1327, "C": 656.2725}
    for glass, nd in ((materials.Glass.n_bk7(), 1.5168), (materials.Glass.f2(), 1.6200)):
        rays = oracle.make_rays(np.array([wl["d"], wl["F"], wl["C"]]), [0, 0, 0], [0, 0, 1])
        n = oracle.material_index_refraction(glass, rays)
        assert abs(n[0] - nd) < 1e-3
        assert n[1] > n[2]


def _read_imd(name):
    data = np.loadtxt(GOLDEN / f"imd_{name}.csv", delimiter=",")
    return data[:, 0], data[:, 1]


IMD_CASES = [
    ("Si", None, materials.Layer("Si"), False),
    ("SiO2", materials.Layer("SiO2", thickness=5.0), materials.Layer("Si"), False),
    ("SiO2_100A", materials.Layer("SiO2", thickness=10.0), materials.Layer("Si"), False),
    ("SiC_Cr", [materials.Layer("SiC", thickness=25.0), materials.Layer("Cr", thickness=5.0)],
     materials.Layer("SiO2"), True),
]


@pytest.mark.parametrize("name,layers,substrate,is_mirror", IMD_CASES, ids=[c[0] for c in IMD_CASES])
def test_multilayer_efficiency_vs_imd(name, layers, substrate, is_mirror):
    """optika/materials/_tests/test_multilayers.py:178-287: IMD curves, rtol 1e-4."""
    w_angstrom, eff_file = _read_imd(name)
    R, T = oracle.multilayer_efficiency(w_angstrom / 10, 1.0, 1.0, layers, substrate)
    eff = (R[0] + R[1]) / 2 if is_mirror else (T[0] + T[1]) / 2
    assert np.allclose(eff, eff_file, rtol=1e-4)


def test_multilayer_bounds_oblique():
    """optika/materials/_tests/test_multilayers.py:41-177: 0 <= R, T <= 1 and
    R + T <= 1 for an absorbing stack at oblique incidence."""
    layers = materials.PeriodicLayerSequence(
        [materials.Layer("Si", thickness=4.02), materials.Layer("Mo", thickness=2.68)], 40)
    wl = np.linspace(55, 70, 31)
    for cos in (1.0, 0.9, 0.5):
        R, T = oracle.multilayer_efficiency(wl, cos, 1.0, layers, materials.Layer("SiO2"))
        for r, t in zip(R, T):
            assert np.all((r >= 0) & (r <= 1)) and np.all((t >= 0) & (t <= 1))
            assert np.all(r + t <= 1 + 1e-12)


def test_matrix_power_matches_repeated_product():
    rng = np.random.default_rng(3)
    m = [[rng.normal(size=4) + 1j * rng.normal(size=4) for _ in range(2)] for _ in range(2)]
    for e in (0, 1, 2, 3, 7, 39):
        p = oracle.trace._power2(m, e, (4,))
        q = oracle.trace._identity2((4,))
        for _ in range(e):
            q = oracle.trace._matmul2(q, m)
        for i in range(2):
            for j in range(2):
                assert np.allclose(p[i][j], q[i][j], rtol=1e-11)


# -- rulings ------------------------------------------------------------------------


RULINGS = [
    rulings.Rulings(spacing=1e-3, diffraction_order=1),
    rulings.MeasuredRulings(spacing=1e-3, diffraction_order=1,
                            efficiency_measured=(np.linspace(300, 700, 9),
                                                 np.exp(-np.square(np.linspace(300, 700, 9) / 1000) / 2))),
    rulings.SinusoidalRulings(spacing=1e-3, depth=10.0, diffraction_order=[0, 1, 2]),
    rulings.SquareRulings(spacing=1e-3, depth=10.0, diffraction_order=[0, 1, 2]),
    rulings.SawtoothRulings(spacing=1e-3, depth=10.0, diffraction_order=[0, 1, 2]),
    rulings.TriangularRulings(spacing=1e-3, depth=10.0, diffraction_order=[0, 1, 2]),
    rulings.RectangularRulings(spacing=1e-3, depth=10.0, ratio_duty=0.3, diffraction_order=[0, 1, 2]),
]


@pytest.mark.parametrize("rl", RULINGS, ids=[type(r).__name__ for r in RULINGS])
def test_rulings_efficiency_in_unit_interval(rl):
    """optika/rulings/_rulings_test.py:53-75 with the reference's instances."""
    normal = np.array([[0.0], [0.0], [-1.0]])
    for rays in ref_rays():
        for order in oracle.trace._orders_of(rl):
            e = oracle.rulings_efficiency(rl, order, rays, normal)
            assert np.all(e >= 0) and np.all(e <= 1)


def test_holographic_rulings_refocus():
    """The docstring example of HolographicRulingSpacing
    (optika/rulings/_spacing.py:141-219): rays from the first recording point,
    diffracted in order 1 off a mirror, pass through the second point."""
    x1 = np.array([[5.0], [0.0], [-10.0]])
    x2 = np.array([[10.0], [0.0], [-10.0]])
    rl = rulings.Rulings(
        spacing=rulings.HolographicRulingSpacing(x1=(5.0, 0.0, -10.0), x2=(10.0, 0.0, -10.0),
                                                 wavelength=500.0),
        diffraction_order=1)
    pos = np.stack([np.linspace(-5, 5, 5), np.zeros(5), np.zeros(5)])
    normal = np.array([[0.0], [0.0], [-1.0]])
    d_in = pos - x1
    d_in = d_in / np.sqrt((d_in ** 2).sum(0))
    kappa, scale = oracle.spacing_vector(rl, pos, normal)
    a = oracle.incident_effective(np.full(5, 500.0), d_in, 1.0, normal, 1, kappa, scale)
    b = oracle.snells_law(a, 1.0, 1.0, normal, True)
    to_x2 = x2 - pos
    cross = oracle.trace._cross(b, to_x2)
    assert np.allclose(cross, 0, atol=1e-9)


def test_spacing_positive_length():
    """optika/rulings/_spacing_test.py:20-40."""
    p = np.stack([np.linspace(-1, 1, 6), np.linspace(-1, 1, 6), np.zeros(6)])
    normal = np.array([[0.0], [0.0], [-1.0]])
    for sp in [rulings.ConstantRulingSpacing(constant=1e-3, normal=(1, 0, 0)),
               rulings.Polynomial1dRulingSpacing(coefficients={0: 1e-3, 1: 2.0, 2: 3e3},
                                                 normal=(1, 0, 0),
                                                 transformation=Cartesian3dRotationZ(30.0))]:
        k, _ = oracle.spacing_vector(rulings.Rulings(spacing=sp), p, normal)
        assert np.all(np.sqrt((k ** 2).sum(0)) > 0)


def test_direction_matches_rotations_and_angles_inverts():
    """optika/_util_test.py:20-48."""
    for ax, ay in ((1.0, 2.0), (0.3, -0.7)):
        d = oracle.direction_from_angles(np.array([ax]), np.array([ay]))[:, 0]
        ry = oracle.trace._rotation_matrix("y", -ax)
        rx = oracle.trace._rotation_matrix("x", ay)
        assert np.allclose(d, ry @ rx @ np.array([0.0, 0.0, 1.0]))
        back_x = -np.degrees(np.arctan2(d[0], d[2]))
        back_y = -np.degrees(np.arcsin(d[1] / np.linalg.norm(d)))
        assert np.allclose([back_x, back_y], [ax, ay])


# -- apertures ----------------------------------------------------------------------


def _winding_inside(x, y, vx, vy):
    """Winding number by summing signed angles (independent of crossing rules)."""
    ang = np.zeros_like(x)
    nv = len(vx)
    for i in range(nv):
        j = (i + 1) % nv
        a1 = np.arctan2(vy[i] - y, vx[i] - x)
        a2 = np.arctan2(vy[j] - y, vx[j] - x)
        d = a2 - a1
        d = (d + np.pi) % (2 * np.pi) - np.pi
        ang += d
    return np.abs(ang) > np.pi


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_point_in_polygon_matches_winding_number(seed):
    rng = np.random.default_rng(seed)
    nv = rng.integers(5, 12)
    th = np.sort(rng.uniform(0, 2 * np.pi, nv))
    r = rng.uniform(0.3, 1.0, nv)  # star-shaped, generally non-convex
    vx, vy = r * np.cos(th), r * np.sin(th)
    x, y = rng.uniform(-1.1, 1.1, 4000), rng.uniform(-1.1, 1.1, 4000)
    a = oracle.point_in_polygon(x, y, vx, vy)
    b = _winding_inside(x, y, vx, vy)
    assert np.array_equal(a, b)


def test_aperture_edges_inclusive_and_flags():
    """optika/apertures/_apertures.py:309-312, 962: <= comparisons, inverted and
    active flags."""
    p = np.array([[10.0, 10.0, 0.0, 10.000001], [0.0, 0.0, 5.0, 0.0], [0.0, 0.0, 0.0, 0.0]])
    circ = apertures.CircularAperture(10.0)
    assert list(oracle.aperture_mask(circ, p)) == [True, True, True, False]
    rect = apertures.RectangularAperture(half_width=(10.0, 5.0))
    assert list(oracle.aperture_mask(rect, p)) == [True, True, True, False]
    inv = apertures.CircularAperture(10.0, inverted=True)
    assert list(oracle.aperture_mask(inv, p)) == [False, False, False, True]
    off = apertures.CircularAperture(1.0, active=False)
    assert np.all(oracle.aperture_mask(off, p))


def test_regular_polygon_and_rectangle_vertices_agree():
    """A rectangle tested as a polygon of its reference vertices gives the
    rectangle mask away from the edges."""
    rect = apertures.RectangularAperture(half_width=(3.0, 2.0))
    rng = np.random.default_rng(5)
    p = np.stack([rng.uniform(-4, 4, 2000), rng.uniform(-3, 3, 2000), np.zeros(2000)])
    v = rect.vertices
    a = oracle.point_in_polygon(p[0], p[1], np.asarray(v.x), np.asarray(v.y))
    b = oracle.aperture_mask(rect, p)
    far = (np.abs(np.abs(p[0]) - 3) > 1e-9) & (np.abs(np.abs(p[1]) - 2) > 1e-9)
    assert np.array_equal(a[far], b[far])


# -- systems --------------------------------------------------------------------


def test_transformation_convention_newtonian_docstring():
    """optika/systems/_sequential.py:1853-1912: with [RotationY(135), Translation]
    applied in order, light reflected by the fold mirror travels to the sensor
    at x = -50 mm."""
    primary_z, fold_z, sensor_x = 200.0, 50.0, 50.0
    primary = surfaces.Surface(
        sag=sags.ParabolicSag(focal_length=-(primary_z - fold_z + sensor_x)),
        material=materials.Mirror(),
        transformation=Cartesian3dTranslation(z=primary_z))
    fold = surfaces.Surface(
        material=materials.Mirror(),
        transformation=TransformationList([Cartesian3dRotationY(135.0),
                                           Cartesian3dTranslation(z=fold_z)]))
    sensor = sensors.ImagingSensor(
        width_pixel=20e-3, num_pixel=Cartesian2dVectorArray(128, 128),
        transformation=TransformationList([Cartesian3dRotationY(-90.0),
                                           Cartesian3dTranslation(x=-sensor_x, z=fold_z)]))
    system = systems.SequentialSystem(surfaces=[primary, fold], sensor=sensor)
    px = np.array([5.0, -5.0, 0.0, 10.0])
    py = np.array([0.0, 5.0, -10.0, 0.0])
    z = np.zeros(4)
    world = oracle.raytrace(system, 500.0, z, z, px, py, accumulate=True)[0]
    hit = world[-1]["position"]
    assert np.allclose(hit[0], -sensor_x) and np.allclose(hit[2], fold_z)
    local = oracle.rayfunction(system, 500.0, z, z, px, py)[0]["position"]
    # the on-axis parabola focuses every parallel ray onto the sensor origin
    assert np.allclose(local[:2], 0, atol=1e-9)
    assert np.all(oracle.rayfunction(system, 500.0, z, z, px, py)[0]["unvignetted"])


def test_collect_matches_bruteforce_histogram():
    """Sensor binning: np.histogramdd rules (last edge inclusive, outside
    dropped, vignetted rays carry no flux) checked by an explicit loop."""
    sensor = sensors.ImagingSensor(width_pixel=1.0, num_pixel=Cartesian2dVectorArray(4, 3))
    rng = np.random.default_rng(4)
    n = 500
    x = rng.uniform(-2.5, 2.5, n)
    y = rng.uniform(-2.0, 2.0, n)
    x[:5] = [2.0, -2.0, 0.0, 1.0, 2.0000001]
    y[:5] = [1.5, -1.5, 0.5, -0.5, 0.0]
    rays = oracle.make_rays(rng.uniform(500, 600, n), np.stack([x, y, np.zeros(n)]),
                            [0.0, 0.0, 1.0], intensity=rng.uniform(0, 1, n),
                            unvignetted=rng.uniform(0, 1, n) > 0.2)
    edges = np.array([500.0, 550.0, 600.0])
    image, _ = oracle.collect(sensor, rays, edges)
    brute = np.zeros((2, 4, 3))
    ex = np.linspace(-2, 2, 5)
    ey = np.linspace(-1.5, 1.5, 4)
    for i in range(n):
        if not rays["unvignetted"][i]:
            continue
        w = rays["wavelength"][i]
        bw = 0 if w < 550 else 1
        if not (ex[0] <= x[i] <= ex[-1] and ey[0] <= y[i] <= ey[-1]):
            continue
        bx = min(int(np.searchsorted(ex, x[i], side="right")) - 1, 3)
        by = min(int(np.searchsorted(ey, y[i], side="right")) - 1, 2)
        brute[bw, bx, by] += rays["intensity"][i]
    assert np.allclose(image, brute, rtol=1e-13, atol=0)


def test_newtonian_on_axis_focus_is_perfect():
    """A paraboloid images an on-axis point at infinity perfectly: every
    unvignetted on-axis ray lands on the sensor origin."""
    from optika_amd import workloads

    g = workloads.newtonian_grid_c1()
    out = oracle.rayfunction(workloads.newtonian(), g.wavelength, g.field_x, g.field_y,
                             g.pupil_x, g.pupil_y)[0]
    on = (g.field_x == 0) & out["unvignetted"]
    assert on.sum() > 3000
    assert np.max(np.abs(out["position"][:2, on])) < 1e-9
    d = out["direction"]
    assert np.allclose(np.sqrt((d ** 2).sum(0)), 1)
    # on axis, the r = 100 mm primary is the only stop: a ray passes exactly
    # when its pupil radius is within the aperture
    r = np.hypot(g.pupil_x, g.pupil_y)
    axis = g.field_x == 0
    assert np.array_equal(out["unvignetted"][axis], r[axis] <= 100.0)
