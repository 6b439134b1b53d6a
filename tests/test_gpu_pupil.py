"""Pupil reductions on the device (optk_pupil_reduce behind vignetting,
distortion and area_effective) against numpy reductions of the oracle's
trace of the same grid, and the effective area of the stop-test Newtonian
against its entrance pupil in closed form."""

import numpy as np
import pytest
import torch

from oracle import stops as S
from oracle import trace as T
from optika_amd import pupil, workloads
from optika_amd.vectors import Cartesian2dVectorArray

pytestmark = pytest.mark.gpu


def _grid():
    wl = np.array([500.0, 650.0])
    f = np.linspace(-1, 1, 5)
    fx, fy = [a.ravel() for a in np.meshgrid(f, f, indexing="ij")]
    p = np.linspace(-1, 1, 9)
    px, py = [a.ravel() for a in np.meshgrid(p, p, indexing="ij")]
    return wl, fx, fy, px, py


def _oracle_rays(system, wl, fx, fy, px, py):
    """Oracle rays of the W x F x P grid (pupil fastest), normalized
    coordinates denormalized by the oracle's own stop solve."""
    out = []
    for w in wl:
        FX, FY, PX, PY = np.broadcast_arrays(fx[:, None], fy[:, None], px[None, :], py[None, :])
        a, b, c, d = S.denormalize(system, w, FX.ravel(), FY.ravel(), PX.ravel(), PY.ravel())
        out.append(T.rayfunction(system, w, a, b, c, d)[0])
    return out


def test_reductions_match_numpy_of_oracle_rays():
    print("case: pupil reductions vs oracle", flush=True)
    system = workloads.stop_newtonian()
    wl, fx, fy, px, py = _grid()
    rays, shape = pupil.trace_outer(system, wl, Cartesian2dVectorArray(fx, fy),
                                    Cartesian2dVectorArray(px, py))
    st = pupil.reduce(rays.outputs, shape)
    W, F, P = shape
    ref = _oracle_rays(system, wl, fx, fy, px, py)
    for w in range(W):
        m = ref[w]["unvignetted"].reshape(F, P)
        x = ref[w]["position"][0].reshape(F, P)
        y = ref[w]["position"][1].reshape(F, P)
        got_m = rays.outputs.unvignetted.reshape(W, F, P)[w].cpu().numpy()
        assert (got_m == m).mean() > 0.999
        same = (got_m == m).all(axis=1)
        np.testing.assert_array_equal(st["count"][0, w][same], m.sum(1)[same])
        np.testing.assert_allclose(st["sum_x"][0, w][same], np.where(m, x, 0).sum(1)[same],
                                   rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(st["all_y"][0, w], y.sum(1), rtol=1e-9, atol=1e-9)
    # the same reduction of the device's own rays, in numpy: equal to rounding
    x = rays.outputs.position.x.reshape(W, F, P).cpu().numpy()
    m = rays.outputs.unvignetted.reshape(W, F, P).cpu().numpy()
    np.testing.assert_array_equal(st["count"][0], m.sum(-1))
    np.testing.assert_allclose(st["sum_x"][0], np.where(m, x, 0).sum(-1), rtol=1e-13, atol=1e-12)


def test_models_from_device_reductions():
    print("case: vignetting / distortion models", flush=True)
    system = workloads.stop_newtonian()
    wl, fx, fy, px, py = _grid()
    f = Cartesian2dVectorArray(fx, fy)
    p = Cartesian2dVectorArray(px, py)
    v = system.vignetting(wl, f, p)
    assert np.all(np.isfinite(v.illumination[v.where]))
    np.testing.assert_allclose(v.illumination[v.where].reshape(2, -1).mean(1), 1.0, rtol=1e-12)
    d = system.distortion(wl, f, p, degree=2)
    # a spherical mirror images angle to position linearly (f = 1000 mm) to
    # within its distortion: sensor x ~ -1000 tan(field x)
    # the quadratic fit leaves the third-order distortion: < 1e-4 mm here
    got = d.distort(d.wavelength, Cartesian2dVectorArray(d.field_x, d.field_y)).x
    np.testing.assert_allclose(got, d.sensor_x, atol=1e-4)
    back = d.undistort(d.wavelength, Cartesian2dVectorArray(d.sensor_x, d.sensor_y))
    np.testing.assert_allclose(back.x, d.field_x, atol=1e-5)
    assert np.allclose(d.sensor_x, -1000 * np.tan(np.deg2rad(d.field_x)), rtol=2e-2, atol=1e-3) or \
        np.allclose(d.sensor_x, 1000 * np.tan(np.deg2rad(d.field_x)), rtol=2e-2, atol=1e-3)


def test_area_effective_is_the_entrance_pupil():
    """The stop-test Newtonian's only clipping aperture is the 10 mm pupil
    stop, imaged by the f = 1000 mm primary to an entrance pupil of radius
    10 / 0.75 mm: with efficiency 1 everywhere the effective area is its
    area, pi (40/3)^2 = 558.5 mm^2, to the grid's sampling error."""
    print("case: area_effective", flush=True)
    system = workloads.stop_newtonian()
    f = Cartesian2dVectorArray(np.array([0.0, 0.5, -0.5]), np.array([0.0, 0.0, 0.5]))
    v = np.linspace(-1, 1, 41)
    m = system.area_effective(wavelength=np.array([500.0, 600.0]), field=f,
                              pupil=Cartesian2dVectorArray(v, v), seed=1)
    want = np.pi * (10.0 / 0.75) ** 2
    area = m([500.0, 600.0])
    assert np.all(np.abs(area - want) < 0.02 * want), area
    assert abs(m(550.0) - area.mean()) < 1e-9
