"""``optk_denormalize``: the normalized grid's map onto the stop rays' extent
(the reference's optika/systems/_sequential.py:748-789, ``_denormalize_grid``)
in one launch.  Its float64 values equal the reference's expression
``ptp * (c + 1) / 2 + lo`` evaluated by numpy, bit for bit; through
``stops.denormalize`` they equal the same expression over the solved stop
extents for per-ray wavelengths and for field and pupil grids of different
shapes."""

import ctypes

import numpy as np
import pytest
import torch

from optika_amd import _lib, _trace, stops
from optika_amd.vectors import Cartesian2dVectorArray
from stop_systems import newtonian

pytestmark = pytest.mark.gpu


def _numpy(c, idx, lo, ptp):
    w = np.zeros(c.shape, np.int64) if idx is None else np.clip(idx, 0, lo.size - 1)
    return ptp[w] * (c + 1) / 2 + lo[w]


@pytest.mark.parametrize("n,ncomp,W,indexed", [(1, 1, 1, False), (1000, 4, 1, False),
                                               (70001, 4, 7, True), (4097, 2, 3, True)])
def test_values_equal_the_reference_expression(n, ncomp, W, indexed):
    print(f"case: denormalize n={n} ncomp={ncomp} W={W} indexed={indexed}", flush=True)
    rng = np.random.default_rng(n + ncomp)
    c = rng.uniform(-1, 1, (ncomp, n))
    lo = rng.normal(size=(ncomp, W))
    ptp = rng.uniform(0, 3, (ncomp, W))
    # out-of-range entries are clamped into the table
    idx = rng.integers(-2, W + 2, n) if indexed else None
    dev = torch.device("cuda")
    cin = torch.as_tensor(c, device=dev)
    lo_t, ptp_t = torch.as_tensor(lo, device=dev), torch.as_tensor(ptp, device=dev)
    out = torch.empty_like(cin)
    a = _lib.DenormalizeArgs()
    for k in range(ncomp):
        a.in_[k], a.out[k] = cin[k].data_ptr(), out[k].data_ptr()
        a.lo[k], a.ptp[k] = lo_t[k].data_ptr(), ptp_t[k].data_ptr()
    idx_t = torch.as_tensor(idx, device=dev) if indexed else None
    a.table_index = idx_t.data_ptr() if indexed else None
    a.n_table = W
    _lib.check(_lib.lib().optk_denormalize(ctypes.byref(a), n, _trace.stream_handle()))
    got = out.cpu().numpy()
    for k in range(ncomp):
        assert np.array_equal(got[k], _numpy(c[k], idx, lo[k], ptp[k])), k
    # in place: out = in
    for k in range(ncomp):
        a.out[k] = cin[k].data_ptr()
    _lib.check(_lib.lib().optk_denormalize(ctypes.byref(a), n, _trace.stream_handle()))
    assert np.array_equal(cin.cpu().numpy(), got)


def _extents(system, wavelengths):
    rays = stops.rayfunction_stops(system, np.asarray(wavelengths, np.float64), 21, 21)
    fld, pup = stops.field_pupil(system, rays)
    ext = []
    for v in (fld, pup):
        lo = v.min(axis=(1, 2))
        ext.append((lo, v.max(axis=(1, 2)) - lo))
    return ext


def test_per_ray_wavelengths():
    print("case: denormalize per-ray wavelengths", flush=True)
    system = newtonian()
    rng = np.random.default_rng(3)
    n = 5000
    wl = rng.choice([600.0, 450.0, 500.0], n)
    u = rng.uniform(-1, 1, (4, n))
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
    f, p = stops.denormalize(system, t(wl), Cartesian2dVectorArray(t(u[0]), t(u[1])),
                             Cartesian2dVectorArray(t(u[2]), t(u[3])))
    uniq, inv = np.unique(wl, return_inverse=True)
    (flo, fptp), (plo, pptp) = _extents(system, uniq)
    want = [_numpy(u[0], inv, flo[0], fptp[0]), _numpy(u[1], inv, flo[1], fptp[1]),
            _numpy(u[2], inv, plo[0], pptp[0]), _numpy(u[3], inv, plo[1], pptp[1])]
    for g, w in zip((f.x, f.y, p.x, p.y), want):
        assert np.array_equal(g.cpu().numpy(), w)


def test_field_and_pupil_of_different_shapes():
    print("case: denormalize broadcast shapes", flush=True)
    system = newtonian()
    dev = torch.device("cuda")
    fx = torch.linspace(-1, 1, 5, dtype=torch.float64, device=dev)[:, None]
    fy = torch.linspace(-1, 1, 3, dtype=torch.float64, device=dev)[None, :]
    px = torch.linspace(-1, 1, 7, dtype=torch.float64, device=dev)
    f, p = stops.denormalize(system, 500.0, Cartesian2dVectorArray(fx, fy),
                             Cartesian2dVectorArray(px, 0.25), normalized_pupil=True)
    assert tuple(f.x.shape) == (5, 1) and tuple(f.y.shape) == (1, 3)
    assert tuple(p.x.shape) == (7,) and tuple(p.y.shape) == ()
    (flo, fptp), (plo, pptp) = _extents(system, [500.0])
    assert np.array_equal(f.x.cpu().numpy(), _numpy(fx.cpu().numpy(), None, flo[0], fptp[0]))
    assert np.array_equal(f.y.cpu().numpy(), _numpy(fy.cpu().numpy(), None, flo[1], fptp[1]))
    assert np.array_equal(p.x.cpu().numpy(), _numpy(px.cpu().numpy(), None, plo[0], pptp[0]))
    assert np.array_equal(p.y.cpu().numpy(), _numpy(np.float64(0.25), None, plo[1], pptp[1]))
