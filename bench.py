"""Benchmark: rays·surfaces/s of the sequential trace on one MI355X.

Workload (one "step"): ``SequentialSystem.rayfunction`` of the Newtonian
telescope (primary, fold mirror, sensor: 3 counted surfaces) over 2^24
rays with pupil positions uniform on the 100 mm disc, field angles uniform
in +-0.25 deg and wavelengths uniform in 400-700 nm (seed 1).  Inputs are
resident in HBM before the timed region; every step traces all rays through
every surface into preallocated output buffers.

Prints one JSON line (see the fields at the end of ``main``).  ``value`` is
the float64 trace -- the drop-in form, identical to the oracle to 1e-12.
The roofline object is for the dominant kernel (trace_grid_f64), timed with
HIP events on the stream it runs on; ``cpu_baseline`` times the numpy
float64 oracle (a restatement of the reference's vectorized numpy trace) on
a bounded sample on this machine's host.
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
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    fx, fy, px, py, wl = (t(grid.field_x), t(grid.field_y), t(grid.pupil_x), t(grid.pupil_y),
                          t(grid.wavelength))
    plan = _trace.GridTracePlan(system, n, sensor_frame=True)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        plan(fx, fy, px, py, wl)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        plan(fx, fy, px, py, wl)
        ends[k].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    kernel_avg_ms = sum(kernel_ms) / len(kernel_ms)
    value = n * surfaces_counted / (ms_per_step / 1e3)
    print(f"[bench] f64 trace: {ms_per_step:.3f} ms/step, kernel {kernel_avg_ms:.3f} ms",
          file=sys.stderr, flush=True)

    achieved = BYTES_PER_RAY_F64 * n / (kernel_avg_ms / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": TRAFFIC_BYTES_F64_2P24 if n == 1 << 24 else None,
        "kernel": "trace_grid_f64",
        "kernel_avg_ms": round(kernel_avg_ms, 4),
        "bytes_per_ray": BYTES_PER_RAY_F64,
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

    out = {
        "metric": "rays·surfaces/s on 1 MI355X: FP64 vs FP32 vs surrogate; RMS spot error vs FP64",
        "value": value,
        "unit": "rays·surfaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "newtonian telescope rayfunction, 2^24 rays (pupil uniform on r=100 mm "
                        "disc, field +-0.25 deg, 400-700 nm, seed 1)",
            "rays": n,
            "surfaces": surfaces_counted,
            "form": "float64 (drop-in, oracle parity 1e-12)",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
