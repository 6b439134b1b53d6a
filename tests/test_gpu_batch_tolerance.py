"""The batched tolerance analysis at its configured size: 4096 perturbed
spectrographs (decenters sigma 50 um, tilts sigma 0.01 deg, radii sigma
0.1 %, line-spacing coefficients sigma 0.5 %), ray counts log-uniform in
[1000, 64000) per instance (61.7M rays, both diffraction orders), in one
persistent launch per form.

* Oracle parity (float64): eight instances -- the smallest and the largest
  ray count, the most tilted and most decentred primary and grating, the
  largest line-spacing error, the largest radius error -- traced by the
  CPU oracle from their own (tilted, decentred, rescaled, VLS-scaled)
  surfaces equal the batch's rays to 1e-12 with identical masks
  (the reference's n-dimensional configurations,
  optika/systems/_sequential.py:2125-2132, optika/surfaces.py:386-395,
  here one instance at a time).
* Statistics: for every instance and path, the kernel's count, sums, RMS
  spot radius and throughput equal reductions of the returned rays
  (torch float64 segment sums on the device, 1e-9 relative).
* Float32 form: the same batch; per-instance spot RMS within 1e-4 mm of the
  float64 form's and counts equal to within the rays at an aperture edge.
"""

import numpy as np
import pytest
import torch

import oracle
from optika_amd import BatchedSystem, workloads

from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    print("case: tolerance batch, 4096 instances (building)", flush=True)
    systems, counts, grid = workloads.tolerance_batch(4096)
    assert len(systems) == 4096 and 55e6 < counts.sum() < 70e6
    return systems, counts, grid


def _inputs(grid, dtype):
    f = lambda a: torch.as_tensor(np.asarray(a), device="cuda").to(dtype).contiguous()  # noqa: E731
    return [f(grid.field_x), f(grid.field_y), f(grid.pupil_x), f(grid.pupil_y), f(grid.wavelength)]


def _segment_stats(out, counts, paths):
    """Per (instance, path): count, sum I, sum x, sum y, sum x^2, sum y^2,
    RMS spot radius about the centroid, throughput = sum I / rays, by
    device segment sums of the returned rays."""
    n = int(counts.sum())
    inst = torch.repeat_interleave(torch.arange(len(counts), device="cuda"),
                                   torch.as_tensor(counts, device="cuda"))
    res = torch.zeros((len(counts), paths, 8), dtype=torch.float64, device="cuda")
    for p in range(paths):
        sl = slice(p * n, (p + 1) * n)
        m = out.t["unvignetted"][sl].double()
        x = out.t["position"][0][sl].double()
        y = out.t["position"][1][sl].double()
        w = out.t["intensity"][sl].double()
        x, y, w = torch.where(m > 0, x, 0.0), torch.where(m > 0, y, 0.0), torch.where(m > 0, w, 0.0)
        cols = [m, w, x, y, x * x, y * y]
        for k, c in enumerate(cols):
            res[:, p, k].index_add_(0, inst, c)
        c = res[:, p, 0]
        mx, my = res[:, p, 2] / c, res[:, p, 3] / c
        var = res[:, p, 4] / c - mx * mx + res[:, p, 5] / c - my * my
        res[:, p, 6] = torch.where(c > 0, torch.sqrt(torch.clamp(var, min=0.0)),
                                   torch.full_like(c, float("nan")))
        res[:, p, 7] = res[:, p, 1] / torch.as_tensor(counts, device="cuda", dtype=torch.float64)
    return res


def _pick(systems, counts):
    """Eight instances that span the perturbations."""
    pert = [s.perturbation for s in systems]
    mag = lambda v: float(np.hypot(*v))  # noqa: E731
    keys = [
        int(np.argmin(counts)), int(np.argmax(counts)),
        int(np.argmax([mag(p.primary_tilt) for p in pert])),
        int(np.argmax([mag(p.primary_decenter) for p in pert])),
        int(np.argmax([mag(p.grating_tilt) for p in pert])),
        int(np.argmax([mag(p.grating_decenter) for p in pert])),
        int(np.argmax([max(abs(v) for v in p.spacing) for p in pert])),
        int(np.argmax([abs(p.primary_radius) + max(abs(v) for v in p.grating_radii) for p in pert])),
    ]
    return list(dict.fromkeys(keys))


def test_tolerance_batch_full_size(c4):
    systems, counts, grid = c4
    offsets = np.concatenate([[0], np.cumsum(counts)])
    n = int(offsets[-1])
    b = BatchedSystem(systems)
    P = b.num_paths
    spots = {}
    for precision, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        print(f"case: tolerance batch full size {precision}, {n} rays x {P} paths", flush=True)
        ins = _inputs(grid, dtype)
        stats, out = b.trace(*ins, counts=counts, precision=precision, outputs=True, graph=False)
        torch.cuda.synchronize()
        got = stats.cpu().numpy()
        want = _segment_stats(out, counts, P).cpu().numpy()
        ok = np.isclose(got, want, rtol=1e-9, atol=1e-9, equal_nan=True)
        assert ok.all(), f"{precision}: stats differ at {np.argwhere(~ok)[:5]}"
        spots[precision] = got
        if precision == "f64":
            for i in _pick(systems, counts):
                lo, hi = int(offsets[i]), int(offsets[i + 1])
                print(f"instance {i}: {hi - lo} rays vs oracle", flush=True)
                ref = oracle.rayfunction(systems[i], grid.wavelength[lo:hi], grid.field_x[lo:hi],
                                         grid.field_y[lo:hi], grid.pupil_x[lo:hi],
                                         grid.pupil_y[lo:hi])
                assert len(ref) == P
                for p in range(P):
                    sl = slice(p * n + lo, p * n + hi)
                    g = {
                        "position": np.stack([t[sl].cpu().numpy() for t in out.t["position"]]),
                        "direction": np.stack([t[sl].cpu().numpy() for t in out.t["direction"]]),
                        "unvignetted": out.t["unvignetted"][sl].cpu().numpy(),
                    }
                    for f in ("wavelength", "intensity", "attenuation", "index_refraction"):
                        g[f] = out.t[f][sl].cpu().numpy()
                    assert_parity(g, ref[p], f"instance {i} path {p}")
        del out, stats, ins
        torch.cuda.empty_cache()
    # float32 form against float64: spot RMS and counts per instance
    c64, c32 = spots["f64"][..., 0], spots["f32"][..., 0]
    assert np.abs(c64 - c32).max() <= 0.001 * counts.max()
    s64, s32 = spots["f64"][..., 6], spots["f32"][..., 6]
    fin = np.isfinite(s64) & np.isfinite(s32) & (c64 > 100)
    assert fin.sum() > 1000
    assert np.abs(s64[fin] - s32[fin]).max() < 1e-4
