"""The oracle's stop solve (oracle/stops.py) against what fixes it:

* the reference's own known answers -- ``field_max`` of its Newtonian and
  grazing-incidence test systems equals the object aperture's half-angle
  within 1e-6 degrees (optika/systems/_sequential_test.py:591-598, 665-672);
* the Newtonian system's entrance pupil from paraxial optics: the 10 mm stop
  250 mm in front of the f = 1000 mm primary images to a virtual pupil of
  radius 10 / 0.75 mm, 1000 / 3 mm behind the mirror, so the marginal ray of
  the edge field meets the object plane at 13.33 + 833.3 tan(0.05 deg) mm;
* the grazing system's pupil: rays from the edge field meeting the
  paraboloid's rim, in closed form;
* the solved rays traced forward land on the pupil stop's outline.
CPU only."""

import numpy as np
import pytest

from oracle import stops as S
from oracle import trace as T
from stop_systems import RADIUS_FIELD_GRAZING, RADIUS_FIELD_NEWTONIAN, grazing, newtonian


@pytest.mark.parametrize("make,radius", [(newtonian, RADIUS_FIELD_NEWTONIAN),
                                         (grazing, RADIUS_FIELD_GRAZING)])
def test_field_max_matches_object_aperture(make, radius):
    (flo, fhi), (plo, phi) = S.field_pupil_extent(make(), [500.0])
    assert np.all(np.abs(fhi[:, 0] - radius) < 1e-6)
    assert np.all(np.abs(flo[:, 0] + radius) < 1e-6)
    assert np.all(phi > plo)


def test_newtonian_entrance_pupil_paraxial():
    (flo, fhi), (plo, phi) = S.field_pupil_extent(newtonian(), [500.0])
    want = 10.0 / 0.75 + (500.0 + 1000.0 / 3.0) * np.tan(np.deg2rad(RADIUS_FIELD_NEWTONIAN))
    # third-order aberrations of the f/75 beam stay inside 0.1 %
    assert abs(phi[0, 0] - want) < 1e-3 * want and abs(plo[0, 0] + want) < 1e-3 * want


def test_grazing_pupil_is_the_paraboloid_rim():
    """The pupil stop is the paraboloid itself: the extreme ray leaves the
    object plane and meets the rim (r = 260 mm, sag r^2 / 4f = -8.45 mm) at
    the edge-field direction, an exact closed form."""
    (_, _), (plo, phi) = S.field_pupil_extent(grazing(), [500.0])
    r = np.sin(np.deg2rad(RADIUS_FIELD_GRAZING))
    want = 260.0 + (2500.0 - 260.0 ** 2 / 8000.0) * r / np.sqrt(1 - r * r)
    assert abs(phi[0, 0] - want) < 1e-8 and abs(plo[0, 0] + want) < 1e-8


def test_stop_rays_land_on_the_pupil_outline():
    sys_ = newtonian()
    r = S.stop_rays(sys_, [500.0], 21, 21)
    surf = [s for s, _, _ in S.surfaces_all(sys_)]
    rays = T.make_rays(500.0, r["position"].reshape(3, -1), r["direction"].reshape(3, -1))
    out = T.propagate(surf[1:3], rays)[0]
    local = T.transform_inverse(surf[2].transformation, out["position"])[0]
    radius = np.hypot(local[0], local[1])
    assert np.all(np.abs(radius - 10.0) < 1e-8)
    # the Newton residual: each ray meets its own outline point
    want = S.wire(surf[2].aperture, 21)[:2]
    got = local[:2].reshape(2, 21, 21)
    assert np.all(np.abs(got - want[:, None, :]) < 1e-8)


def test_wires_follow_the_reference_sampling():
    from optika_amd import apertures as A

    w = S.wire(A.RectangularAperture(half_width=(2.0, 1.0)), 9)
    np.testing.assert_allclose(w[:2, [0, 2, 4, 6, 8]], [[2, -2, -2, 2, 2], [1, 1, -1, -1, 1]],
                               atol=1e-15)
    c = S.wire(A.CircularAperture(radius=3.0), 5)
    np.testing.assert_allclose(c[:2], [[3, 0, -3, 0, 3], [0, 3, 0, -3, 0]], atol=1e-14)
    s = S.wire(A.CircularSectorAperture(radius=1.0, angle_start=0.0, angle_stop=90.0), 10)
    assert s.shape == (3, 10) and np.allclose(s[:2, 0], 0) and np.allclose(s[:2, -1], 0)
    # the package's outlines are the same points
    for ap, n in [(A.RectangularAperture(half_width=(2.0, 1.0)), 9), (A.CircularAperture(radius=3.0), 5),
                  (A.CircularSectorAperture(radius=1.0, angle_start=0.0, angle_stop=90.0), 10),
                  (A.OctagonalAperture(radius=2.0), 21), (A.EllipticalAperture(radius=(1.0, 2.0)), 7)]:
        o = S.wire(ap, n)
        p = ap.wire(n)
        np.testing.assert_allclose(np.stack([p.x, p.y, p.z]), o, atol=1e-14)


def test_angles_invert_the_direction_map():
    """``angles`` (optika/_util.py:76-97) undoes ``optika.direction``
    (d = (-cos y sin x, -sin y, cos y cos x)): round trips of random field
    angles, and a hand value (10 deg about y alone gives (-sin 10, 0, cos 10))."""
    rng = np.random.default_rng(5)
    ax, ay = rng.uniform(-30, 30, 200), rng.uniform(-30, 30, 200)
    d = np.stack(T.direction_from_angles(ax, ay))
    back = S.angles(d)
    np.testing.assert_allclose(back[0], ax, atol=1e-12)
    np.testing.assert_allclose(back[1], ay, atol=1e-12)
    t = np.deg2rad(10.0)
    np.testing.assert_allclose(S.angles(np.array([[-np.sin(t)], [0.0], [np.cos(t)]]))[:, 0], [10.0, 0.0],
                               atol=1e-12)
    # a non-unit direction: only the y angle's arcsine divides by the length
    np.testing.assert_allclose(S.angles(3.0 * d)[1], ay, atol=1e-12)


def test_num_paths_counts_the_orders_between_the_stops():
    from optika_amd import workloads

    assert S.num_paths(newtonian()) == 1
    # the spectrograph's grating (orders -1, +1) sits between its pupil stop
    # (the primary) and its field stop
    assert S.num_paths(workloads.spectrograph()) == 2


def test_rayfunction_stops_walks_back_to_the_object_plane():
    """The stop rays walked back to the object: on the object plane, unit
    directions turned into the system (+z for the Newtonian, whose object
    sits in front of the primary), every outline point's ray present."""
    r = S.rayfunction_stops(newtonian(), [500.0], 21, 21)
    p, d = r["position"], r["direction"]
    assert p.shape[0] == 3 and d.shape == p.shape
    assert np.all(np.abs(p[2]) < 1e-9)
    np.testing.assert_allclose(np.sqrt((d * d).sum(0)), 1.0, atol=1e-12)
    assert np.all(d[2] > 0)


def test_denormalize_maps_the_unit_square_onto_the_extent():
    """``denormalize`` (_sequential.py:777-785): -1 and +1 land on the
    extent's ends, 0 on its middle, and the map is linear between."""
    sys_ = newtonian()
    (flo, fhi), (plo, phi) = S.field_pupil_extent(sys_, [500.0])
    u = np.array([-1.0, 0.0, 0.5, 1.0])
    fx, fy, px, py = S.denormalize(sys_, 500.0, u, u, u, u)
    for got, lo, hi in ((fx, flo[0, 0], fhi[0, 0]), (fy, flo[1, 0], fhi[1, 0]),
                        (px, plo[0, 0], phi[0, 0]), (py, plo[1, 0], phi[1, 0])):
        np.testing.assert_allclose(got, [lo, (lo + hi) / 2, lo + 0.75 * (hi - lo), hi], rtol=1e-14,
                                   atol=1e-14)
