"""The kernels' per-ray code, compiled for the host (tests/harness), against
the oracle: bit-identical geometry, identical masks, multilayer efficiency
to 1e-13.  CPU only; the same sources build the gfx950 kernels."""

import numpy as np
import pytest

import oracle
from harness import multilayer, trace_grid
from optika_amd import materials, workloads


def _assert_trace_equal(h, o, exact=True):
    assert len(h) == len(o)
    for hh, oo in zip(h, o):
        assert np.array_equal(hh["unvignetted"], oo["unvignetted"])
        for k in ("position", "direction", "wavelength"):
            if exact:
                np.testing.assert_array_equal(hh[k], oo[k])
            else:
                np.testing.assert_allclose(hh[k], oo[k], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(hh["intensity"], oo["intensity"], rtol=1e-13, atol=1e-15)


def test_newtonian_c1_bit_identical():
    g = workloads.newtonian_grid_c1()
    s = workloads.newtonian()
    _assert_trace_equal(trace_grid(s, g),
                        oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y))


def test_newtonian_random_bit_identical():
    g = workloads.newtonian_grid_random(20000, seed=11)
    s = workloads.newtonian()
    _assert_trace_equal(trace_grid(s, g),
                        oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y))


def test_spectrograph_both_orders():
    g = workloads.spectrograph_grid(4000, seed=12)
    s = workloads.spectrograph()
    _assert_trace_equal(trace_grid(s, g),
                        oracle.rayfunction(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y))


def test_global_frame_trace():
    g = workloads.newtonian_grid_random(3000, seed=13)
    s = workloads.newtonian()
    h = trace_grid(s, g, sensor_frame=False)
    o = oracle.raytrace(s, g.wavelength, g.field_x, g.field_y, g.pupil_x, g.pupil_y)
    _assert_trace_equal(h, o)


@pytest.mark.parametrize("cos", [1.0, 0.95, 0.7])
def test_multilayer_kernel_code(cos):
    layers = materials.PeriodicLayerSequence(
        [materials.Layer("Si", thickness=4.02), materials.Layer("Mo", thickness=2.68)], 40)
    sub = materials.Layer("SiO2")
    wl = np.linspace(12.0, 70.0, 97)
    Rs, Rp, Ts, Tp = multilayer(layers, sub, wl, cos, 1.0 + 0j)
    (oRs, oRp), (oTs, oTp) = oracle.multilayer_efficiency(wl, cos, 1.0 + 0j, layers, sub)
    for a, b in ((Rs, oRs), (Rp, oRp), (Ts, oTs), (Tp, oTp)):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)


def _tables_of(system):
    from optika_amd import _compile

    rec = _compile.build_records(system.surfaces_all)
    t = np.asarray(rec.tables)
    return [t[rec.chemicals[i].offset:rec.chemicals[i].offset + rec.chemicals[i].count]
            for i in range(rec.num_chemicals)]


@pytest.mark.parametrize("kind", ["measured", "linear", "short", "nonpositive"])
def test_chemical_grid_brackets(kind):
    """The n/k index grid gives numpy.searchsorted's bracket for every
    wavelength (knots, midpoints, float32-rounding neighbours, random) and
    starts less than one knot from it on average in the EUV band of the
    measured log-spaced tables, so a lookup is one load and a short walk."""
    from harness import grid_bracket

    rng = np.random.default_rng(4)
    if kind == "measured":
        tables = _tables_of(workloads.spectrograph())
        assert min(t.size for t in tables) > 1000 and tables[0][-1] / tables[0][0] > 1e6
    elif kind == "linear":
        tables = [np.linspace(500.0, 700.0, 301)]
    elif kind == "short":
        tables = [np.array([100.0, 400.0]), np.array([1.0, 2.0, 2.5])]
    else:
        tables = [np.linspace(-5.0, 5.0, 11)]
    for w in tables:
        mids = 0.5 * (w[1:] + w[:-1])
        x = np.concatenate([w[:-1], mids, np.nextafter(w[1:-1], -np.inf), np.nextafter(w[:-1], np.inf),
                            w[:-1].astype(np.float32).astype(np.float64),
                            np.exp(rng.uniform(np.log(max(w[0], 1e-3)), np.log(w[-1]), 20000))
                            if w[0] > 0 else rng.uniform(w[0], w[-1], 20000)])
        x = x[(x >= w[0]) & (x < w[-1])]
        j, steps, cells = grid_bracket(w, x)
        if kind == "nonpositive":
            assert cells == 0  # no grid: the binary search brackets instead
            continue
        want = np.clip(np.searchsorted(w, x, side="right") - 1, 0, w.size - 2)
        np.testing.assert_array_equal(j, want)
        assert cells <= 16384
        if kind == "measured":
            eband = (x > 100.0) & (x < 2000.0)  # EUV .. soft x-ray, where traces look up
            assert steps.max() <= 16 and steps[eband].mean() < 1.0, (steps.max(), steps[eband].mean())


def test_small_angle_sincos_against_numpy():
    """The float64 kernels' series sin and cos of field angles (|x| <= 0.01
    rad) against numpy's: within 1 ulp over the range, exact at 0, odd and
    even in x; and the direction it gives through a whole trace stays
    within the drop-in's 1e-12 (checked against the oracle on the GPU)."""
    from harness import small_angle_sincos

    x = np.concatenate([np.linspace(-0.01, 0.01, 200_001), [0.0, 1e-300, -1e-12, 0.01, -0.01]])
    s, c = small_angle_sincos(x)
    ulp_s = np.spacing(np.abs(np.sin(x)))
    ulp_c = np.spacing(np.abs(np.cos(x)))
    assert np.all(np.abs(s - np.sin(x)) <= ulp_s)
    assert np.all(np.abs(c - np.cos(x)) <= ulp_c)
    assert s[-5] == 0.0 and c[-5] == 1.0
    s2, c2 = small_angle_sincos(-x)
    assert np.array_equal(s2, -s) and np.array_equal(c2, c)
