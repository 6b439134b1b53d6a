"""Batched persistent trace (optk_batch_*) of perturbed spectrographs with
ragged ray counts.

* float64 form: every instance's rays equal the single-system float64 trace
  of that instance (same per-ray code): positions/directions within 1e-12
  relative, identical masks.
* float32 form: every instance's rays within 1e-5 mm / 1e-7 of the
  single-system float32 trace (the kernels are compiled separately, so FMA
  contraction may differ in the last bits).
* statistics: count, sums, RMS spot radius and throughput equal numpy
  reductions of the returned rays (1e-9 relative).
* hipGraph replay: repeated replays, with other device work between them,
  reproduce the eager statistics.
"""

import numpy as np
import pytest
import torch

from optika_amd import STAT_FIELDS, BatchedSystem, workloads
from optika_amd.vectors import Cartesian2dVectorArray

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_case():
    systems, counts, grid = workloads.tolerance_batch(48, seed=11, min_rays=200, max_rays=6000)
    counts[5] = 0  # an instance without rays
    n = int(counts.sum())
    grid = grid.subset(n)
    return systems, counts, grid


def _inputs(grid, dtype):
    f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device="cuda").to(dtype)  # noqa: E731
    return [f(grid.field_x), f(grid.field_y), f(grid.pupil_x), f(grid.pupil_y), f(grid.wavelength)]


def _single(system, ins, lo, hi, precision):
    fx, fy, px, py, wl = (t[lo:hi] for t in ins)
    return system.rayfunction(wavelength=wl, field=Cartesian2dVectorArray(fx, fy),
                              pupil=Cartesian2dVectorArray(px, py), normalized_field=False,
                              normalized_pupil=False, precision=precision).outputs


def _stats_numpy(out, offsets, paths):
    n = offsets[-1]
    x = out.t["position"][0].double().cpu().numpy().reshape(paths, n)
    y = out.t["position"][1].double().cpu().numpy().reshape(paths, n)
    w = out.t["intensity"].double().cpu().numpy().reshape(paths, n)
    m = out.t["unvignetted"].cpu().numpy().reshape(paths, n)
    res = np.zeros((len(offsets) - 1, paths, 8))
    for i in range(len(offsets) - 1):
        sl = slice(offsets[i], offsets[i + 1])
        for p in range(paths):
            mm = m[p, sl]
            xs, ys, ws = x[p, sl][mm], y[p, sl][mm], w[p, sl][mm]
            c = mm.sum()
            res[i, p, :6] = [c, ws.sum(), xs.sum(), ys.sum(), (xs * xs).sum(), (ys * ys).sum()]
            res[i, p, 6] = np.sqrt(((xs - xs.mean()) ** 2 + (ys - ys.mean()) ** 2).mean()) if c else np.nan
            res[i, p, 7] = ws.sum() / (offsets[i + 1] - offsets[i]) if offsets[i + 1] > offsets[i] else 0.0
    return res


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_matches_single_traces(batch_case, precision):
    systems, counts, grid = batch_case
    dtype = torch.float64 if precision == "f64" else torch.float32
    print(f"case: batch {precision}, {len(systems)} instances, {int(counts.sum())} rays", flush=True)
    ins = _inputs(grid, dtype)
    b = BatchedSystem(systems)
    stats, out = b.trace(*ins, counts=counts, precision=precision, outputs=True, graph=False)
    torch.cuda.synchronize()
    offsets = np.concatenate([[0], np.cumsum(counts)])
    n = int(offsets[-1])
    P = b.num_paths
    assert P == 2
    for i in range(len(systems)):
        lo, hi = int(offsets[i]), int(offsets[i + 1])
        if hi == lo:
            continue
        ref = _single(systems[i], ins, lo, hi, precision)
        for p in range(P):
            r = ref[p]
            got_m = out.t["unvignetted"][p * n + lo:p * n + hi]
            assert torch.equal(got_m, r.unvignetted), f"instance {i} path {p}: masks"
            pos, dirs = list(r.position), list(r.direction)
            for k in range(3):
                a = out.t["position"][k][p * n + lo:p * n + hi].double()
                e = pos[k].double()
                fin = torch.isfinite(e)
                if precision == "f64":
                    tol = 1e-12 * torch.clamp(torch.abs(e[fin]), min=1.0)
                else:
                    tol = 1e-5 + 1e-7 * torch.abs(e[fin])
                assert torch.all(torch.abs(a[fin] - e[fin]) <= tol), f"instance {i} path {p} pos {k}"
                a = out.t["direction"][k][p * n + lo:p * n + hi].double()
                e = dirs[k].double()
                fin = torch.isfinite(e)
                tol = 1e-12 if precision == "f64" else 1e-7
                assert torch.all(torch.abs(a[fin] - e[fin]) <= tol), f"instance {i} path {p} dir {k}"
    want = _stats_numpy(out, offsets, P)
    got = stats.cpu().numpy()
    assert got.shape == (len(systems), P, len(STAT_FIELDS))
    ok = np.isclose(got, want, rtol=1e-9, atol=1e-9, equal_nan=True)
    assert ok.all(), f"stats differ at {np.argwhere(~ok)[:5]}"
    # the empty instance: no rays, no throughput, undefined spot
    assert got[5, :, 0].sum() == 0 and np.isnan(got[5, 0, 6]) and got[5, 0, 7] == 0.0


def test_batch_graph_replays_and_autotune(batch_case):
    systems, counts, grid = batch_case
    print("case: batch graph replay + autotune", flush=True)
    ins = _inputs(grid, torch.float32)
    b = BatchedSystem(systems)
    eager = b.trace(*ins, counts=counts, precision="f32", graph=False).clone()
    torch.cuda.synchronize()
    stats = b.stats_tensor()
    for rep in range(4):
        b.trace(*ins, counts=counts, precision="f32", graph=True, stats=stats)
        torch.cuda.synchronize()
        # other device work between replays
        junk = torch.randn(1 << 20, device="cuda").sum().item()
        assert np.isfinite(junk)
        got = stats.cpu().numpy()
        ok = np.isclose(got, eager.cpu().numpy(), rtol=1e-12, atol=1e-12, equal_nan=True)
        assert ok.all(), f"replay {rep} differs"
    best = b.autotune(*ins, counts=counts, precision="f32")
    assert best.block in (64, 128, 256, 512) and 1 <= best.rays_per_thread <= 8 and best.ms > 0
    tuned = b.trace(*ins, counts=counts, precision="f32", graph=True)
    torch.cuda.synchronize()
    ok = np.isclose(tuned.cpu().numpy(), eager.cpu().numpy(), rtol=1e-10, atol=1e-10, equal_nan=True)
    assert ok.all()


def test_batch_rejects_bad_counts(batch_case):
    systems, counts, grid = batch_case
    ins = _inputs(grid, torch.float32)
    b = BatchedSystem(systems[:4])
    with pytest.raises(ValueError):
        b.trace(*ins, counts=counts[:4], precision="f32")


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_specialised_kernel_equals_generic(batch_case, precision):
    """Every instance of the tolerance batch keeps the tilted-spectrograph
    kinds, so the batch runs the specialised persistent kernel; the generic
    one (OPTK_TRACE_GENERIC) gives the same rays: identical in float64, to
    float32 rounding in float32."""
    from optika_amd import _trace

    systems, counts, grid = batch_case
    dtype = torch.float64 if precision == "f64" else torch.float32
    print(f"case: batch specialised vs generic {precision}", flush=True)
    sigs = {_trace.GridTracePlan(s, 1, precision="f64").cs.signature for s in systems[:4]}
    assert sigs == {_trace.SIGNATURE_NAMES.index("tilted spectrograph")}
    ins = _inputs(grid, dtype)
    b = BatchedSystem(systems)
    s_spec, spec = b.trace(*ins, counts=counts, precision=precision, outputs=True, graph=False)
    s_gen, gen = b.trace(*ins, counts=counts, precision=precision, outputs=True, graph=False,
                         generic=True)
    torch.cuda.synchronize()
    assert torch.equal(spec.t["unvignetted"], gen.t["unvignetted"])
    for name in ("position", "direction"):
        for k in range(3):
            a, e = spec.t[name][k].double(), gen.t[name][k].double()
            fin = torch.isfinite(e)
            assert torch.equal(torch.isfinite(a), fin), f"{name}[{k}] finiteness"
            if precision == "f64":
                assert torch.equal(a[fin], e[fin]), f"{name}[{k}]"
            else:
                tol = (1e-5 if name == "position" else 1e-7) + 1e-7 * torch.abs(e[fin])
                assert torch.all(torch.abs(a[fin] - e[fin]) <= tol), f"{name}[{k}]"
    a, e = s_spec.cpu().numpy(), s_gen.cpu().numpy()
    assert np.array_equal(a[..., 0], e[..., 0]), "ray counts"
    if precision == "f64":  # the sums are accumulated with atomics: order may differ
        ok = np.isclose(a, e, rtol=1e-9, atol=1e-9, equal_nan=True)
        assert ok.all(), f"stats differ at {np.argwhere(~ok)[:5]}"


def test_mixed_layout_batch_runs_generic_and_matches_single_traces():
    """A batch whose instances do not share one surface-kind signature (the
    nominal spectrograph, whose primary is only decentred, beside tilted
    ones, whose primaries take a full affine map) runs the generic kernel:
    every instance still equals its single-system float64 trace."""
    from optika_amd import _trace

    tilted = workloads.Perturbation(primary_tilt=(0.02, -0.01), grating_tilt=(0.01, 0.0),
                                    primary_decenter=(0.05, 0.0))
    systems = [workloads.spectrograph(), workloads.spectrograph(tilted), workloads.spectrograph()]
    sigs = [_trace.GridTracePlan(s, 1, precision="f64").cs.signature for s in systems]
    assert len(set(sigs)) == 2, sigs
    counts = np.array([3000, 1500, 2200])
    n = int(counts.sum())
    grid = workloads.spectrograph_grid(n, seed=12)
    print("case: mixed-layout batch f64", flush=True)
    ins = _inputs(grid, torch.float64)
    b = BatchedSystem(systems)
    _, out = b.trace(*ins, counts=counts, precision="f64", outputs=True, graph=False)
    torch.cuda.synchronize()
    offsets = np.concatenate([[0], np.cumsum(counts)])
    for i in range(len(systems)):
        lo, hi = int(offsets[i]), int(offsets[i + 1])
        ref = _single(systems[i], ins, lo, hi, "f64")
        for p in range(b.num_paths):
            r = ref[p]
            assert torch.equal(out.t["unvignetted"][p * n + lo:p * n + hi], r.unvignetted)
            for k, e in enumerate(r.position):
                a = out.t["position"][k][p * n + lo:p * n + hi]
                fin = torch.isfinite(e)
                tol = 1e-12 * torch.clamp(torch.abs(e[fin]), min=1.0)
                assert torch.all(torch.abs(a[fin] - e[fin]) <= tol), f"instance {i} path {p} pos {k}"
