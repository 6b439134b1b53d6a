"""Benchmark: rays·surfaces/s of the sequential trace on one MI355X.

Workload (one "step"): ``SequentialSystem.rayfunction`` of the Newtonian
telescope (primary, fold mirror, sensor: 3 counted surfaces) over 2^24
rays with pupil positions uniform on the 100 mm disc, field angles uniform
in +-0.25 deg and wavelengths uniform in 400-700 nm (seed 1).  Inputs are
resident in HBM before the timed region; every step traces all rays through
every surface into preallocated output buffers.

Prints one JSON line (see the fields at the end of ``main``).  ``value`` is
the float32 compensated form; ``forms`` holds the float64 drop-in form
(identical to the oracle to 1e-12) and the float32 form's RMS spot and
direction error against it over all rays of the run.  The roofline object
is for the float32 kernel (trace_grid_f32), timed with HIP events on the
stream it runs on; ``cpu_baseline`` times the numpy float64 oracle (a
restatement of the reference's vectorized numpy trace) on a bounded sample
on this machine's host.
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# algorithmic bytes per ray of the grid-input rayfunction (float64):
#   in:  field x, y, pupil x, y, wavelength           5 x 8 = 40 B
#   out: position 3, direction 3, intensity           7 x 8 = 56 B
#        unvignetted mask                                      1 B
# (wavelength, attenuation and index are invariant through an all-mirror
#  system and are returned as the input / constants, not rewritten)
BYTES_PER_RAY_F64 = 97

# HBM bytes per launch of trace_grid_f64 at 2^24 rays, from rocprofv3 --pmc:
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a 256 MiB device copy);
# the runs and commands are in profiles/trace_grid_f64_newtonian_2p24.md
TRAFFIC_BYTES_F64_2P24 = (342198 * 2.0 + 950396 * 1.0) * 1024

# float32 form: in 5 x 4 B, out 7 x 4 B + 1 B mask
BYTES_PER_RAY_F32 = 49
TRAFFIC_BYTES_F32_2P24 = None  # not yet measured with rocprofv3 --pmc

PIXEL_MM = 0.015


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--rays", type=int, default=1 << 24)
    p.add_argument("--cpu-rays", type=int, default=1 << 18,
                   help="rays in the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a 256 MiB device copy (to scale HBM byte counters)")
    return p.parse_args()


def cpu_baseline(system, grid, sample: int) -> dict:
    """Time the numpy float64 oracle on a bounded sample: 1 warm-up, median
    of 3 repeats, one host thread."""
    import oracle

    sub = grid.subset(sample)
    times = []
    for rep in range(4):
        t0 = time.perf_counter()
        oracle.rayfunction(system, sub.wavelength, sub.field_x, sub.field_y, sub.pupil_x,
                           sub.pupil_y)
        t1 = time.perf_counter()
        if rep > 0:
            times.append(t1 - t0)
        print(f"[bench] cpu baseline repeat {rep}: {t1 - t0:.2f} s", file=sys.stderr, flush=True)
    med = statistics.median(times)
    return {
        "value": sample * 3 / med,
        "unit": "rays·surfaces/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{sample} rays of the same workload (first {sample} of the seeded grid), "
                  f"numpy float64 oracle, median of 3 after 1 warm-up, "
                  f"{os.cpu_count()} host CPUs present, 1 thread used",
    }


def _time_form(plan, inputs, steps, warmup, stream):
    """Warm up, then time `steps` launches: wall clock between two device
    synchronisations, and each launch with HIP events on `stream`."""
    import torch

    for _ in range(warmup):
        plan(*inputs)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        plan(*inputs)
        ends[k].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kernel_ms = [s_.elapsed_time(e_) for s_, e_ in zip(starts, ends)]
    return (t1 - t0) * 1e3 / steps, sum(kernel_ms) / len(kernel_ms)


def _spot_error(r64, r32) -> dict:
    """float32 form against float64 form, over rays unvignetted in both."""
    import torch

    m64, m32 = r64.unvignetted, r32.unvignetted
    both = m64 & m32
    dx = (r32.position.x.double() - r64.position.x)[both] / PIXEL_MM
    dy = (r32.position.y.double() - r64.position.y)[both] / PIXEL_MM
    err = torch.sqrt(dx * dx + dy * dy)
    dd = torch.stack([(a.double() - b)[both] for a, b in zip(r32.direction, r64.direction)])
    derr = torch.sqrt((dd * dd).sum(0))
    return {
        "rms_spot_px": torch.sqrt((err * err).mean()).item(),
        "max_spot_px": err.max().item(),
        "rms_direction": torch.sqrt((derr * derr).mean()).item(),
        "mask_agreement": (m64 == m32).double().mean().item(),
        "rays_compared": int(both.sum().item()),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(rays shard with no exchange; run one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch

    from optika_amd import _trace, workloads

    dev = torch.device("cuda")
    n = args.rays
    system = workloads.newtonian()
    surfaces_counted = workloads.NEWTONIAN_SURFACES
    print(f"[bench] generating {n} rays", file=sys.stderr, flush=True)
    grid = workloads.newtonian_grid_random(n, seed=1)
    # float32 inputs; the float64 form traces the same values widened
    t32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)  # noqa: E731
    in32 = [t32(grid.field_x), t32(grid.field_y), t32(grid.pupil_x), t32(grid.pupil_y),
            t32(grid.wavelength)]
    in64 = [t.double() for t in in32]
    plan64 = _trace.GridTracePlan(system, n, sensor_frame=True, precision="f64")
    plan32 = _trace.GridTracePlan(system, n, sensor_frame=True, precision="f32")
    stream = torch.cuda.current_stream()

    ms64, k64 = _time_form(plan64, in64, args.steps, args.warmup, stream)
    print(f"[bench] f64 trace: {ms64:.3f} ms/step, kernel {k64:.3f} ms", file=sys.stderr, flush=True)
    ms32, k32 = _time_form(plan32, in32, args.steps, args.warmup, stream)
    print(f"[bench] f32 trace: {ms32:.3f} ms/step, kernel {k32:.3f} ms", file=sys.stderr, flush=True)
    err = _spot_error(plan64(*in64), plan32(*in32))
    torch.cuda.synchronize()
    print(f"[bench] f32 vs f64: {err}", file=sys.stderr, flush=True)

    value64 = n * surfaces_counted / (ms64 / 1e3)
    value32 = n * surfaces_counted / (ms32 / 1e3)
    achieved = BYTES_PER_RAY_F32 * n / (k32 / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": TRAFFIC_BYTES_F32_2P24 if n == 1 << 24 else None,
        "kernel": "trace_grid_f32",
        "kernel_avg_ms": round(k32, 4),
        "bytes_per_ray": BYTES_PER_RAY_F32,
    }

    if args.calibrate_copy:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(system, grid, min(args.cpu_rays, n))

    f64_achieved = BYTES_PER_RAY_F64 * n / (k64 / 1e3) / 1e9
    out = {
        "metric": "rays·surfaces/s on 1 MI355X: FP64 vs FP32 vs surrogate; RMS spot error vs FP64",
        "value": value32,
        "unit": "rays·surfaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "newtonian telescope rayfunction, 2^24 rays (pupil uniform on r=100 mm "
                        "disc, field +-0.25 deg, 400-700 nm, seed 1)",
            "rays": n,
            "surfaces": surfaces_counted,
            "form": "float32 compensated form (float32 I/O); accuracy vs the float64 form "
                    "in forms.f32",
        },
        "forms": {
            "f64": {"value": value64, "ms_per_step": ms64, "kernel_avg_ms": round(k64, 4),
                    "kernel": "trace_grid_f64", "bytes_per_ray": BYTES_PER_RAY_F64,
                    "achieved_gbs": round(f64_achieved, 1),
                    "traffic": TRAFFIC_BYTES_F64_2P24 if n == 1 << 24 else None,
                    "note": "drop-in form: oracle parity 1e-12, identical masks"},
            "f32": {"value": value32, "ms_per_step": ms32, "kernel_avg_ms": round(k32, 4),
                    "speedup_vs_f64": value32 / value64, **err},
            "surrogate": None,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
