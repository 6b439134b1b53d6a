"""Benchmark: rays·surfaces/s of the sequential trace on one MI355X.

Default workload (one "step"): ``SequentialSystem.rayfunction`` of the
Newtonian telescope (primary, fold mirror, sensor: 3 counted surfaces) over
2^24 rays with pupil positions uniform on the 100 mm disc, field angles
uniform in +-0.25 deg and wavelengths uniform in 400-700 nm (seed 1).
Inputs are resident in HBM before the timed region; every step traces all
rays through every surface into preallocated output buffers.

Prints one JSON line.  ``value`` is the float64 drop-in form (the
reference's arithmetic, identical to the oracle to 1e-12); ``forms`` holds
the metric's other legs, each with its error against the float64 form: the
float32 compensated form on the same rays, and the learned surrogate of the
EUV spectrograph (trained here on 2e8 float64-traced rays, timed on 1e9
inference rays, judged against the float64 trace of that system).
``roofline`` is the float64 kernel's (``forms.f32.roofline`` the float32
kernel's), timed with HIP events on the stream it runs on.  ``cpu_baseline``
times the numpy float64 oracle (a restatement of the reference's
vectorized numpy trace) on this machine's host cores, in worker processes
started before anything touches the GPU: N cores (one process per core) and
one core.

Other workloads (``--workload``): ``spectrograph`` (the EUV spectrograph,
4 counted surfaces, two diffraction orders, Mo/Si multilayers on both
mirrors; 50M rays) and ``tolerance`` (4096 perturbed spectrographs with
log-uniform ray counts in one persistent batched launch, autotuned and
replayed from a hipGraph), ``surrogate`` (the learned surrogate alone) and
``stops`` (normalized field/pupil coordinates: the stop solve on the
device, then the trace of the denormalized grid, both inside each step) and
``rays`` (``propagate_rays``, the ray entry point, on the Newtonian).
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
#   in:  field x, y, pupil x, y                       4 x 8 = 32 B
#        wavelength, when the trace depends on it       + 8 B
#   out: position 3, direction 3, intensity           7 x 8 = 56 B
#        unvignetted mask                                      1 B
# (wavelength, attenuation and index are invariant through an all-mirror
#  system and are returned as the input / constants, not rewritten; the
#  Newtonian's specialised kernels do not read the wavelength at all: no
#  surface of it depends on it)
def bytes_per_ray(precision: str, reads_wavelength: bool) -> int:
    w = 8 if precision == "f64" else 4
    return w * (4 + int(reads_wavelength)) + 7 * w + 1


# HBM bytes per launch of the Newtonian's specialised trace_grid kernels at
# 2^24 rays, from rocprofv3 --pmc: reads = 128 B x TCC_EA0_RDREQ_128B +
# 64 B x TCC_EA0_RDREQ_64B + 32 B x TCC_EA0_RDREQ_32B (a 256 MiB device
# copy in the same run reads exactly 256 MiB by this count), writes =
# WRITE_SIZE KiB (the copy writes 256 MiB by it); the runs and commands are
# in the profile summaries named here
TRAFFIC_BYTES_F64_2P24 = 4194304 * 128 + 1833 * 64 + 939217 * 1024
TRAFFIC_SOURCE_F64 = "profiles/trace_grid_f64_newtonian_2p24.md (rocprofv3 --pmc)"
TRAFFIC_BYTES_F32_2P24 = 2097152 * 128 + 1576 * 64 + 479621 * 1024
TRAFFIC_SOURCE_F32 = "profiles/trace_grid_f32_newtonian_2p24.md (rocprofv3 --pmc)"

PIXEL_MM = 0.015


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", choices=("newtonian", "spectrograph", "tolerance", "surrogate", "stops",
                                          "rays"),
                   default="newtonian")
    p.add_argument("--train-rays", type=int, default=200_000_000,
                   help="surrogate: rays traced (float64) for training")
    p.add_argument("--epochs", type=float, default=2.0, help="surrogate: training epochs")
    p.add_argument("--rays", type=int, default=None,
                   help="rays per step (default: 2^24 newtonian, 50M spectrograph)")
    p.add_argument("--instances", type=int, default=4096, help="tolerance: instances")
    p.add_argument("--cpu-rays", type=int, default=None,
                   help="rays in the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-surrogate", action="store_true",
                   help="default workload: skip the surrogate leg")
    p.add_argument("--infer-rays", type=int, default=1_000_000_000,
                   help="surrogate leg: inference rays per step")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a 256 MiB device copy (to scale HBM byte counters)")
    return p.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(system, grid, sample: int, surfaces: int, paths: int = 1) -> dict:
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
        log(f"cpu baseline repeat {rep}: {t1 - t0:.2f} s")
    med = statistics.median(times)
    return {
        "value": sample * paths * surfaces / med,
        "unit": "rays·surfaces/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{sample} rays of the same workload (first {sample} of the seeded grid"
                  f"{', both diffraction orders' if paths > 1 else ''}), numpy float64 oracle, "
                  f"median of 3 after 1 warm-up, {os.cpu_count()} host CPUs present, "
                  f"1 thread used",
    }


# the oracle's job in the worker processes of the N-core baseline (set
# before the pool forks, so every worker inherits the system and the grid)
_CPU_JOB = None


def _cpu_chunk(span):
    import oracle

    system, grid = _CPU_JOB
    sub = grid.subset(span[1], start=span[0])
    t0 = time.perf_counter()
    oracle.rayfunction(system, sub.wavelength, sub.field_x, sub.field_y, sub.pupil_x, sub.pupil_y)
    return time.perf_counter() - t0


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines(system, grid, surfaces: int, rays_1: int, rays_n: int, paths: int = 1) -> dict:
    """The numpy float64 oracle on this host: N cores (``rays_n`` rays split
    over one worker process per core, wall clock of the whole map) and one
    core (``rays_1`` rays in this process); 1 warm-up and the median of 3
    repeats each.  Runs before anything touches the GPU: the workers are
    forked from this process, which must not hold a GPU context.  The
    numpy oracle holds the GIL through its many small array operations, so
    threads do not scale it; processes do."""
    import multiprocessing

    global _CPU_JOB
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))  # the box's CPU share for one GPU
    one = cpu_baseline(system, grid, rays_1, surfaces, paths)
    _CPU_JOB = (system, grid)
    per = (rays_n + cores - 1) // cores
    spans = [(k * per, min(per, rays_n - k * per)) for k in range(cores) if k * per < rays_n]
    walls = []
    with multiprocessing.get_context("fork").Pool(len(spans)) as pool:
        for rep in range(4):
            t0 = time.perf_counter()
            pool.map(_cpu_chunk, spans, chunksize=1)
            dt = time.perf_counter() - t0
            log(f"cpu baseline {len(spans)} processes, repeat {rep}: {dt:.2f} s")
            if rep > 0:
                walls.append(dt)
    _CPU_JOB = None
    med = statistics.median(walls)
    return {
        "value": rays_n * paths * surfaces / med,
        "unit": "rays·surfaces/s",
        "cores": len(spans),
        "kind": "port",
        "sample": f"first {rays_n} rays of the same seeded workload"
                  f"{' (both diffraction orders)' if paths > 1 else ''}, split over "
                  f"{len(spans)} worker processes (one per core), numpy float64 oracle, "
                  f"median wall clock of 3 after 1 warm-up",
        "cpu_model": _cpu_model(),
        "host_cpus": os.cpu_count(),
        "one_core": {"value": one["value"], "cores": 1, "sample": one["sample"]},
    }


def time_calls(fn, steps, warmup, stream):
    """Warm up, then time `steps` calls: wall clock between two device
    synchronisations, and each call with HIP events on `stream`."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        fn()
        ends[k].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kernel_ms = [s_.elapsed_time(e_) for s_, e_ in zip(starts, ends)]
    return (t1 - t0) * 1e3 / steps, sum(kernel_ms) / len(kernel_ms)


def spot_error(r64, r32) -> dict:
    """float32 form against float64 form, over rays unvignetted in both."""
    import torch

    m64, m32 = r64.unvignetted, r32.unvignetted
    both = m64 & m32
    dx = (r32.position.x.double() - r64.position.x)[both] / PIXEL_MM
    dy = (r32.position.y.double() - r64.position.y)[both] / PIXEL_MM
    err = torch.sqrt(dx * dx + dy * dy)
    dd = torch.stack([(a.double() - b)[both] for a, b in zip(r32.direction, r64.direction)])
    derr = torch.sqrt((dd * dd).sum(0))
    out = {
        "rms_spot_px": torch.sqrt((err * err).mean()).item() if err.numel() else None,
        "max_spot_px": err.max().item() if err.numel() else None,
        "rms_direction": torch.sqrt((derr * derr).mean()).item() if derr.numel() else None,
        "mask_agreement": (m64 == m32).double().mean().item(),
        "rays_compared": int(both.sum().item()),
    }
    if isinstance(r64.intensity, torch.Tensor) and err.numel():
        i64 = r64.intensity[both]
        i32 = r32.intensity[both].double()
        rel = torch.abs(i32 - i64) / torch.clamp(torch.abs(i64), min=1e-30)
        out["rms_intensity_rel"] = torch.sqrt((rel * rel).mean()).item()
    return out


def header(args, metric_value, ms, dtype, config, forms, roofline, cpu) -> dict:
    return {
        "metric": "rays·surfaces/s on 1 MI355X: FP64 vs FP32 vs surrogate; RMS spot error vs FP64",
        "value": metric_value,
        "unit": "rays·surfaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
        "forms": forms,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def calibrate_copy(dev):
    import torch

    src = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def two_forms(args, system, grid, n, surfaces, paths, traffic32=None,
              traffic64=None, sensor=None):
    """Time the float64 and float32 forms of a grid trace on the same
    (float32-representable) inputs and compare them over every ray."""
    import torch

    from optika_amd import _trace

    dev = torch.device("cuda")
    t32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)  # noqa: E731
    in32 = [t32(grid.field_x), t32(grid.field_y), t32(grid.pupil_x), t32(grid.pupil_y),
            t32(grid.wavelength)]
    in64 = [t.double() for t in in32]
    plan64 = _trace.GridTracePlan(system, n, sensor_frame=True, precision="f64")
    plan32 = _trace.GridTracePlan(system, n, sensor_frame=True, precision="f32")
    kernel64, kernel32 = plan64.kernel, plan32.kernel
    stream = torch.cuda.current_stream()
    ms64, k64 = time_calls(lambda: plan64(*in64), args.steps, args.warmup, stream)
    log(f"f64 trace: {ms64:.3f} ms/step, kernel {k64:.3f} ms")
    ms32, k32 = time_calls(lambda: plan32(*in32), args.steps, args.warmup, stream)
    log(f"f32 trace: {ms32:.3f} ms/step, kernel {k32:.3f} ms")
    r64, r32 = plan64(*in64), plan32(*in32)
    errs = []
    for p in range(paths):
        a = r64 if paths == 1 else r64[p]
        b = r32 if paths == 1 else r32[p]
        errs.append(spot_error(a, b))
    torch.cuda.synchronize()
    log(f"f32 vs f64: {errs}")
    hist = None
    if sensor is not None:
        from optika_amd import _collect

        # the sensor image (optk_collect) of both forms' rays
        edges, half, num = sensor
        args_h = (edges, (-half[0], half[0]), (-half[1], half[1]), num[0], num[1])
        img64, _ = _collect.histogram(r64, *args_h)
        ms_h, _ = time_calls(lambda: _collect.histogram(r32, *args_h), 5, 1, stream)
        img32, _ = _collect.histogram(r32, *args_h)
        d = (img32 - img64).abs().sum().item()
        # exposure of the binned image (Poisson shot noise + read noise)
        ms_e, _ = time_calls(lambda: _collect.expose(img64, 1.0, 2.0, True, 7), 5, 1, stream)
        e = _collect.expose(img64, 1.0, 0.0, True, 7)
        hist = {"pixels": f"{num[0]}x{num[1]}", "ms_per_image": ms_h,
                "f64_total": img64.sum().item(), "f32_total": img32.sum().item(),
                "l1_difference_rel": d / max(img64.abs().sum().item(), 1e-300),
                "expose_ms_per_image": ms_e,
                "expose_electrons_vs_photons": e.sum().item() / max(img64.sum().item(), 1e-300)}
        log(f"histogram: {hist}")
    units = n * paths * surfaces
    value64 = units / (ms64 / 1e3)
    value32 = units / (ms32 / 1e3)

    def roof(bytes_per_ray, k_ms, kernel, traffic):
        achieved = bytes_per_ray * n * paths / (k_ms / 1e3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic[0] if traffic else None,
                "traffic_source": traffic[1] if traffic else None,
                "kernel": kernel, "kernel_avg_ms": round(k_ms, 4),
                "bytes_per_ray": bytes_per_ray}

    roof64 = roof(bytes_per_ray("f64", plan64.reads_wavelength), k64, kernel64, traffic64)
    roof32 = roof(bytes_per_ray("f32", plan32.reads_wavelength), k32, kernel32, traffic32)
    err = errs[0] if paths == 1 else {f"path{p}": e for p, e in enumerate(errs)}
    forms = {
        "f64": {"value": value64, "ms_per_step": ms64, "kernel_avg_ms": round(k64, 4),
                "kernel": kernel64, "note": "drop-in form: oracle parity 1e-12, identical masks; "
                                            "this is the line's value"},
        "f32": {"value": value32, "ms_per_step": ms32, "kernel_avg_ms": round(k32, 4),
                "kernel": kernel32, "speedup_vs_f64": value32 / value64, "error_vs_f64": err,
                "roofline": roof32},
        "surrogate": None,
    }
    if hist is not None:
        forms["sensor_histogram"] = hist
    return value64, ms64, forms, roof64


def run_newtonian(args):
    from optika_amd import workloads

    n = args.rays or (1 << 24)
    system = workloads.newtonian()
    S = workloads.NEWTONIAN_SURFACES
    log(f"generating {n} rays")
    grid = workloads.newtonian_grid_random(n, seed=1)
    is_2p24 = n == 1 << 24
    # the CPU baselines first: their worker processes fork from this one,
    # which has not touched the GPU yet
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(system, grid, S, rays_1=min(args.cpu_rays or 1 << 22, n),
                            rays_n=min(4 * (args.cpu_rays or 1 << 22), n))
    value, ms, forms, roofline = two_forms(
        args, system, grid, n, S, 1,
        (TRAFFIC_BYTES_F32_2P24, TRAFFIC_SOURCE_F32) if is_2p24 else None,
        (TRAFFIC_BYTES_F64_2P24, TRAFFIC_SOURCE_F64) if is_2p24 else None,
        sensor=([400.0, 700.0], (15.36, 15.36), (2048, 2048)))
    if args.calibrate_copy:
        import torch

        calibrate_copy(torch.device("cuda"))
    if not args.no_surrogate:
        forms["surrogate"] = surrogate_leg(args, steps=max(2, args.steps // 4))[0]
    config = {
        "workload": "newtonian telescope rayfunction, 2^24 rays (pupil uniform on r=100 mm "
                    "disc, field +-0.25 deg, 400-700 nm, seed 1)" if is_2p24 else
                    f"newtonian telescope rayfunction, {n} rays",
        "rays": n,
        "surfaces": S,
        "form": "float64 drop-in form (the reference's arithmetic); the float32 form in "
                "forms.f32, the surrogate (of the EUV spectrograph) in forms.surrogate",
    }
    return header(args, value, ms, "f64", config, forms, roofline, cpu)


def run_spectrograph(args):
    from optika_amd import workloads

    n = args.rays or 50_000_000
    system = workloads.spectrograph()
    S = workloads.SPECTROGRAPH_SURFACES
    log(f"generating {n} rays")
    grid = workloads.spectrograph_grid(n, seed=2)
    value, ms, forms, roofline = two_forms(args, system, grid, n, S, 2)
    cpu = None if args.no_cpu else cpu_baseline(system, grid, min(args.cpu_rays or 1 << 12, n),
                                                S, paths=2)
    config = {
        "workload": f"EUV spectrograph rayfunction, {n} rays x 2 diffraction orders (pupil "
                    "r=40 mm disc, field +-0.05 deg, 58-63 nm, seed 2), Mo/Si 40-period "
                    "multilayers on primary and grating",
        "rays": n,
        "paths": 2,
        "surfaces": S,
        "form": "float64 drop-in form; the float32 form in forms.f32",
    }
    return header(args, value, ms, "f64", config, forms, roofline, cpu)


def run_tolerance(args):
    import torch

    from optika_amd import BatchedSystem, workloads

    N = args.instances
    log(f"building {N} perturbed spectrographs")
    systems, counts, grid = workloads.tolerance_batch(N, seed=3)
    n = int(counts.sum())
    S = workloads.SPECTROGRAPH_SURFACES
    P = 2
    dev = torch.device("cuda")
    t32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)  # noqa: E731
    in32 = [t32(grid.field_x), t32(grid.field_y), t32(grid.pupil_x), t32(grid.pupil_y),
            t32(grid.wavelength)]
    in64 = [t.double() for t in in32]
    b = BatchedSystem(systems)
    log(f"{N} instances, {n} rays; autotuning")
    tune32 = b.autotune(*in32, counts=counts, precision="f32")
    log(f"f32 tune: block {tune32.block}, rays/thread {tune32.rays_per_thread}, {tune32.ms:.3f} ms")
    tune64 = b.autotune(*in64, counts=counts, precision="f64")
    log(f"f64 tune: block {tune64.block}, rays/thread {tune64.rays_per_thread}, {tune64.ms:.3f} ms")
    stats32 = b.stats_tensor()
    stats64 = b.stats_tensor()
    stream = torch.cuda.current_stream()
    ms32, _ = time_calls(lambda: b.trace(*in32, counts=counts, precision="f32", graph=True,
                                         stats=stats32), args.steps, args.warmup, stream)
    log(f"f32 batch: {ms32:.3f} ms/step")
    ms64, _ = time_calls(lambda: b.trace(*in64, counts=counts, precision="f64", graph=True,
                                         stats=stats64), max(2, args.steps // 4), 1, stream)
    log(f"f64 batch: {ms64:.3f} ms/step")
    s32, s64 = stats32.cpu().numpy(), stats64.cpu().numpy()
    live = s64[:, 0, 0] > 0
    rms_err_um = np.abs(s32[live, 0, 6] - s64[live, 0, 6]) * 1e3
    thr_rel = np.abs(s32[live, 0, 7] - s64[live, 0, 7]) / np.maximum(s64[live, 0, 7], 1e-30)
    units = n * P * S
    value32 = units / (ms32 / 1e3)
    value64 = units / (ms64 / 1e3)
    cpu = None
    if not args.no_cpu:
        import oracle

        k = 0
        sample = 0
        while k < N and sample < (args.cpu_rays or 8000):
            sample += int(counts[k])
            k += 1
        t0 = time.perf_counter()
        off = 0
        for i in range(k):
            sub = grid.subset(int(counts[i]), start=off)
            oracle.rayfunction(systems[i], sub.wavelength, sub.field_x, sub.field_y, sub.pupil_x,
                               sub.pupil_y)
            off += int(counts[i])
        dt = time.perf_counter() - t0
        cpu = {"value": sample * P * S / dt, "unit": "rays·surfaces/s", "cores": 1, "kind": "port",
               "sample": f"first {k} instances ({sample} rays, both orders), numpy float64 "
                         f"oracle, one pass, {os.cpu_count()} host CPUs present, 1 thread used"}
    forms = {
        "f64": {"value": value64, "ms_per_step": ms64,
                "tune": {"block": tune64.block, "rays_per_thread": tune64.rays_per_thread}},
        "f32": {"value": value32, "ms_per_step": ms32, "speedup_vs_f64": value32 / value64,
                "tune": {"block": tune32.block, "rays_per_thread": tune32.rays_per_thread},
                "spot_rms_abs_err_um_max": float(rms_err_um.max()),
                "throughput_rel_err_max": float(thr_rel.max())},
        "surrogate": None,
        "per_instance_spot_rms_um": {
            "median": float(np.median(s64[live, 0, 6]) * 1e3),
            "p95": float(np.percentile(s64[live, 0, 6], 95) * 1e3)},
    }
    config = {
        "workload": f"tolerance batch: {N} perturbed EUV spectrographs (decenter 50 um, tilt "
                    "0.01 deg, radii 0.1 %, line spacing 0.5 %, seed 3), ray counts log-uniform "
                    f"1k-64k ({n} rays) x 2 orders, one persistent launch, hipGraph replay",
        "instances": N,
        "rays": n,
        "paths": P,
        "surfaces": S,
        "form": "float64 drop-in form, autotuned; the float32 form in forms.f32",
    }
    return header(args, value64, ms64, "f64", config, forms, None, cpu)


def device_spectrograph_inputs(n: int, seed: int, dev):
    """Float32 device inputs drawn like ``workloads.spectrograph_grid``
    (pupil uniform on the r = 40 mm disc, field +-0.05 deg, 58-63 nm), by a
    seeded device generator: the inference batch of 1e9 rays is made on the
    GPU instead of the host."""
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    u = lambda: torch.rand(n, generator=g, device=dev, dtype=torch.float32)  # noqa: E731
    r = 40.0 * torch.sqrt(u())
    th = (2 * np.pi) * u()
    px, py = r * torch.cos(th), r * torch.sin(th)
    del r, th
    fx = u() * 0.1 - 0.05
    fy = u() * 0.1 - 0.05
    wl = u() * 5.0 + 58.0
    return [fx, fy, px, py, wl]


def surrogate_leg(args, steps: int):
    """The learned surrogate of the spectrograph's order -1 rayfunction:
    train on float64-traced rays, calibrate, then time fused inference
    over ``args.infer_rays`` rays, with and without the flagged rays traced,
    next to the float64 and float32 traces
    of the same rays, and measure the error against the float64 trace on a
    held-out subsample.  Returns (forms entry, roofline, float64 rate,
    float32 rate)."""
    import torch

    from optika_amd import _trace, surrogate, workloads

    dev = torch.device("cuda")
    system = workloads.spectrograph()
    S = workloads.SPECTROGRAPH_SURFACES
    lo = [-0.05, -0.05, -40.0, -40.0, 58.0]
    hi = [0.05, 0.05, 40.0, 40.0, 63.0]
    sg = surrogate.Surrogate(system, lo, hi, path=0, seed=0)
    t0 = time.perf_counter()
    # training rays drawn on the device (as spectrograph_grid draws them on
    # the host), one seeded chunk at a time
    n_train = sg.prepare(lambda m, seed: device_spectrograph_inputs(m, seed, dev), args.train_rays,
                         chunk=1 << 23, poly_rays=1 << 22)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"traced {args.train_rays} training rays ({n_train} usable) in {t1 - t0:.1f} s")
    losses = sg.train(epochs=args.epochs, batch=65536, lr=2e-3, log_every=500)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    adam = max(1, int(args.epochs * n_train / 65536))
    log(f"trained {adam} Adam steps in {t2 - t1:.1f} s; losses {losses[-3:]}")
    del sg.train_u, sg.train_t

    def inputs(n, seed):
        g = workloads.spectrograph_grid(n, seed=seed)
        t32 = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)  # noqa: E731
        return [t32(g.field_x), t32(g.field_y), t32(g.pupil_x), t32(g.pupil_y), t32(g.wavelength)]

    cal = inputs(1 << 20, 501)
    sg.calibrate(*[t.double() for t in cal])
    # accuracy on a held-out subsample, surrogate + fallback vs float64 trace
    ev = inputs(1 << 20, 502)
    ref = sg.trace64(*[t.double() for t in ev])
    out = sg.infer_traced(*ev, precision="f32")  # flagged rays by the float32 trace
    fb = out["list"][:int(out["count"].item())].long()
    x, y, unv = out["x"].double(), out["y"].double(), out["unvignetted"]
    both = ref["unvignetted"] & unv
    err = torch.sqrt((x - ref["x"]) ** 2 + (y - ref["y"]) ** 2)[both] / PIXEL_MM
    nofb = both & ~out["fallback"]
    err_nofb = torch.sqrt((x - ref["x"]) ** 2 + (y - ref["y"]) ** 2)[nofb] / PIXEL_MM
    acc = {"rms_spot_px": torch.sqrt((err ** 2).mean()).item(),
           "rms_spot_px_surrogate_rays": torch.sqrt((err_nofb ** 2).mean()).item(),
           "max_spot_px_surrogate_rays": err_nofb.max().item(),
           "mask_agreement": (ref["unvignetted"] == unv).double().mean().item(),
           "fallback_fraction": fb.numel() / ev[0].numel(),
           "rays_compared": ev[0].numel()}
    log(f"surrogate accuracy: {acc}")
    del ref, out, x, y, unv
    # throughput: fused inference over n rays, in one call
    n = args.infer_rays
    big = device_spectrograph_inputs(n, 7, dev)  # a seed no training chunk uses (those from 1000)
    res = {"x": torch.empty(n, dtype=torch.float32, device=dev),
           "y": torch.empty(n, dtype=torch.float32, device=dev),
           "intensity": torch.empty(n, dtype=torch.float32, device=dev),
           "unvignetted": torch.empty(n, dtype=torch.bool, device=dev),
           "fallback": torch.empty(n, dtype=torch.bool, device=dev),
           "list": torch.empty(n // 4, dtype=torch.int32, device=dev),
           "count": torch.zeros(1, dtype=torch.int32, device=dev)}
    stream = torch.cuda.current_stream()
    ms_inf, k_sur = time_calls(lambda: sg.infer(*big, out=res), steps, 1, stream)
    fb_count = int(res["count"].item())
    log(f"surrogate inference: {ms_inf:.2f} ms for {n} rays, {fb_count} fall back")
    # the whole job: inference, then the flagged rays traced (float32 form)
    # and their results written over the predictions
    ms_sur, _ = time_calls(lambda: sg.infer_traced(*big, out=res, precision="f32"), steps, 1, stream)
    log(f"surrogate with the fallback rays traced: {ms_sur:.2f} ms for {n} rays")
    sg.release_fallback_plan()
    # float64 and float32 traces of the first 2^24 of the same rays
    m = min(n, 1 << 24)
    sub32 = [t[:m].contiguous() for t in big]
    sub64 = [t.double() for t in sub32]
    del big, res
    plan64 = _trace.GridTracePlan(system, m, sensor_frame=True, precision="f64")
    plan32 = _trace.GridTracePlan(system, m, sensor_frame=True, precision="f32")
    ms64, _ = time_calls(lambda: plan64(*sub64), 3, 1, stream)
    ms32, _ = time_calls(lambda: plan32(*sub32), 5, 1, stream)
    del plan64, plan32, sub64, sub32
    torch.cuda.empty_cache()
    v_sur = n * S / (ms_sur / 1e3)  # surfaces of the emulated path per ray
    v64 = m * 2 * S / (ms64 / 1e3)
    v32 = m * 2 * S / (ms32 / 1e3)
    flops_per_ray = 2 * (5 * 256 + 3 * 256 * 256 + 256 * 3)
    achieved = n * flops_per_ray / (k_sur / 1e3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 1), "peak": 2500.0, "unit": "TFLOP/s",
                "frac": round(achieved / 2500.0, 4), "traffic": None,
                "kernel": "surrogate_infer", "kernel_avg_ms": round(k_sur, 3),
                "flops_per_ray": flops_per_ray}
    entry = {"workload": f"learned surrogate of the EUV spectrograph (order -1): 5 -> 4 x 256 SiLU "
                         f"-> 3 MLP, bf16 MFMA, trained on {args.train_rays} float64-traced rays, "
                         f"inference on {n} rays, the rays it flags traced by the float32 form "
                         f"(their order only) inside the timed region",
             "value": v_sur, "ms_per_step": ms_sur, "rays": n, "fallback_rays": fb_count,
             "inference_only": {"value": n * S / (ms_inf / 1e3), "ms_per_step": ms_inf,
                                "note": "the fused kernel alone: fallback rays listed, not traced"},
             "spectrograph_f64_value": v64, "spectrograph_f32_value": v32,
             "speedup_vs_f64": v_sur / v64, "speedup_vs_f32": v_sur / v32,
             "error_vs_f64": acc, "roofline": roofline,
             "training": {"rays": args.train_rays, "usable": n_train, "adam_steps": adam,
                          "batch": 65536, "trace_s": t1 - t0, "train_s": t2 - t1}}
    return entry, roofline, v64, v32


def run_surrogate(args):
    """The surrogate leg alone (its own line)."""
    from optika_amd import workloads

    S = workloads.SPECTROGRAPH_SURFACES
    entry, roofline, v64, v32 = surrogate_leg(args, steps=max(2, args.steps // 4))
    forms = {"f64": {"value": v64}, "f32": {"value": v32}, "surrogate": entry}
    config = {"workload": entry["workload"], "rays": entry["rays"], "surfaces": S,
              "form": "surrogate (value counts surfaces of the emulated path per ray)"}
    cpu = None
    if not args.no_cpu:
        # the reference's path for these rays: the numpy float64 trace
        cpu = cpu_baseline(workloads.spectrograph(), workloads.spectrograph_grid(1 << 12, seed=502),
                           min(args.cpu_rays or 1 << 12, 1 << 12), S, paths=2)
    return header(args, entry["value"], entry["ms_per_step"], "bf16", config, forms, roofline, cpu)


def run_stops(args):
    """Normalized-coordinate rayfunction of the reference's stop-test
    Newtonian: each step solves the stops on the device (21 x 21 Newton
    rays, every iteration's residual and Jacobian columns in one trace) and
    traces the denormalized grid."""
    import torch

    from optika_amd import workloads
    from optika_amd.vectors import Cartesian2dVectorArray

    n = args.rays or 1 << 22
    system = workloads.stop_newtonian()
    S = 3
    dev = torch.device("cuda")
    rng = np.random.default_rng(6)
    u = torch.as_tensor(rng.uniform(-1, 1, (4, n)), device=dev)
    f = Cartesian2dVectorArray(u[0], u[1])
    pp = Cartesian2dVectorArray(u[2], u[3])
    stream = torch.cuda.current_stream()

    def step():
        return system.rayfunction(wavelength=500.0, field=f, pupil=pp)

    from optika_amd import stops as dstops

    def solve():
        # the solve itself (system.rayfunction_stops caches its result, as
        # the reference's cached_property does)
        return dstops.field_pupil(system, dstops.rayfunction_stops(system, np.array([500.0])))

    waves = np.linspace(400.0, 700.0, 64)

    def solve64():
        return dstops.rayfunction_stops(system, waves)

    ms, _ = time_calls(step, args.steps, args.warmup, stream)
    ms_solve, _ = time_calls(solve, args.steps, args.warmup, stream)
    ms_solve64, _ = time_calls(solve64, max(2, args.steps // 4), 1, stream)
    fmax = system.field_max
    cpu = None
    forms = {"stop_solve": {"ms_per_solve": ms_solve, "solves_per_s": 1e3 / ms_solve,
                            "ms_per_solve_64_wavelengths": ms_solve64,
                            "field_max_deg": [float(fmax.x[0]), float(fmax.y[0])],
                            "field_max_expected_deg": workloads.RADIUS_FIELD_NEWTONIAN}}
    # pupil reductions: the reduction kernel over the step's rays (segments
    # of 1024 pupil samples), and a whole area_effective call
    from optika_amd import pupil as dpupil

    out = step().outputs
    seg_p = 1024
    segs = n // seg_p
    shape = (1, segs, seg_p)
    ms_reduce, _ = time_calls(lambda: dpupil.reduce(out, shape), args.steps, args.warmup, stream)
    vtx = np.linspace(-1.0, 1.0, 41)
    fpts = Cartesian2dVectorArray(np.array([0.0, 0.5, -0.5]), np.array([0.0, 0.0, 0.5]))

    def area():
        return system.area_effective(wavelength=np.array([500.0, 600.0]), field=fpts,
                                     pupil=Cartesian2dVectorArray(vtx, vtx), seed=1)

    ms_area, _ = time_calls(area, max(2, args.steps // 4), 1, stream)
    forms["pupil"] = {
        "reduce_ms": ms_reduce, "reduce_rays": n, "reduce_segments": segs,
        "reduce_algorithmic_gbs": n * 25 / (ms_reduce * 1e-3) / 1e9,
        "area_effective_ms": ms_area,
        "area_effective_mm2": [float(a) for a in area()([500.0, 600.0])],
        "area_effective_expected_mm2": float(np.pi * (10.0 / 0.75) ** 2),
        "note": "reduce_ms includes the host copy of the per-segment results",
    }
    if not args.no_cpu:
        from oracle import stops as ostops

        def cpu_ms(w, reps=3):
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                ostops.field_pupil_extent(system, w)
                times.append(time.perf_counter() - t0)
            return statistics.median(times[1:]) * 1e3

        c1 = cpu_ms([500.0])
        c64 = cpu_ms(list(waves))
        forms["stop_solve"]["cpu_oracle_ms_per_solve"] = c1
        forms["stop_solve"]["cpu_oracle_ms_64_wavelengths"] = c64
        cpu = {"value": 1e3 / c64, "unit": "stop solves/s (64 wavelengths)",
               "cores": 1, "kind": "port",
               "sample": "stop solves of the same system at 64 wavelengths (21 x 21 x 64 rays, "
                         "numpy float64 oracle trace), median of 2 after 1 warm-up, 1 thread"}
    config = {"workload": f"normalized rayfunction of the stop-test Newtonian (angular "
                          f"object field stop 0.05 deg, pupil stop 250 mm after the primary), "
                          f"{n} rays uniform in [-1, 1]^4, stop solve + trace per step",
              "rays": n, "surfaces": S, "form": "float64 drop-in form"}
    return header(args, n * S / (ms * 1e-3), ms, "f64", config, forms, None, cpu)


# algorithmic bytes per ray of the ray entry point (optk_trace: position 3,
# direction 3, wavelength, intensity, attenuation, index and the mask in;
# the same ten fields and the mask out)
BYTES_PER_RAY_RAYS_F32 = 10 * 4 + 1 + 10 * 4 + 1
BYTES_PER_RAY_RAYS_F64 = 10 * 8 + 1 + 10 * 8 + 1


def run_rays(args):
    """``propagate_rays`` over the Newtonian's surfaces (the ray entry point,
    ``optk_trace``): 2^24 rays given in the system frame -- on the object
    plane at the seeded grid's pupil positions, directions within ~1e-3 rad
    of the axis -- through every surface, both forms."""
    import torch

    from optika_amd import _trace, workloads
    from optika_amd.rays import RayVectorArray
    from optika_amd.vectors import Cartesian3dVectorArray

    n = args.rays or (1 << 24)
    system = workloads.newtonian()
    S = workloads.NEWTONIAN_SURFACES
    grid = workloads.newtonian_grid_random(n, seed=1)
    surfaces = system.surfaces_all
    cs = _trace.compiled(surfaces)
    stream = torch.cuda.current_stream()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    forms = {}
    for prec, dt in (("f64", torch.float64), ("f32", torch.float32)):
        t = lambda a: torch.as_tensor(np.asarray(a), device="cuda", dtype=dt)  # noqa: E731
        d = torch.randn(3, n, device="cuda", dtype=torch.float64, generator=gen) * 1e-3
        d[2] = 1.0
        d = (d / d.norm(dim=0)).to(dt)
        zero = torch.zeros(n, device="cuda", dtype=dt)
        rays = RayVectorArray(wavelength=t(grid.wavelength),
                              position=Cartesian3dVectorArray(t(grid.pupil_x), t(grid.pupil_y), zero),
                              direction=Cartesian3dVectorArray(d[0], d[1], d[2]),
                              intensity=torch.ones_like(zero), attenuation=zero.clone(),
                              index_refraction=torch.ones_like(zero),
                              unvignetted=torch.ones(n, device="cuda", dtype=torch.bool))
        ms, k = time_calls(lambda: _trace.trace_sequence(surfaces, rays, precision=prec, system=cs),
                           args.steps, args.warmup, stream)
        b = BYTES_PER_RAY_RAYS_F32 if prec == "f32" else BYTES_PER_RAY_RAYS_F64
        forms[prec] = {"value": n * S / (ms * 1e-3), "ms_per_step": ms, "kernel_avg_ms": round(k, 4),
                       "kernel": f"trace_rays_{prec}<0, false>", "bytes_per_ray": b,
                       "achieved_gbs": round(b * n / (k * 1e-3) / 1e9, 1)}
        log(f"{prec} rays: {ms:.3f} ms/step, kernel {k:.3f} ms")
    f32 = forms["f32"]
    roofline = {"bound": "hbm", "achieved": f32["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(f32["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f32["kernel"], "kernel_avg_ms": f32["kernel_avg_ms"],
                "bytes_per_ray": BYTES_PER_RAY_RAYS_F32}
    config = {"workload": f"newtonian telescope propagate_rays (ray entry point), {n} rays in the "
                          "system frame through every surface",
              "rays": n, "surfaces": S, "form": "float32 compensated form; float64 in forms.f64"}
    return header(args, f32["value"], f32["ms_per_step"], "f32", config, forms, roofline, None)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(rays shard with no exchange; run one process per GPU)", file=sys.stderr)
        sys.exit(2)
    run = {"newtonian": run_newtonian, "spectrograph": run_spectrograph,
           "tolerance": run_tolerance, "surrogate": run_surrogate,
           "stops": run_stops, "rays": run_rays}[args.workload]
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
