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
