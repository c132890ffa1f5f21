"""Benchmark of the pulsed-ion-chamber time-stepping loop on one MI355X.

One "step" is one simulation time step over the whole grid: the gather
deposition of that step's tracks (in-pulse steps), the fused Lax-Wendroff
sweep of both species with its scored reduction, the lateral-ring update and
the per-step reduction. The timed region starts with every track of the run
already resident in HBM and ends when the last step's record is written.

Default workload (the mid-size single solve): 56.2 MeV protons, 300 V over a
0.2 cm gap, 540 us pulses at 50 Hz, 8.91 Gy/s, library-default physics, 5 um
voxels over a sampled radius of 0.05 cm (208 x 208 x 406 voxels, 3388 steps),
seed 1, FP32 fields with FP64 recombination and scored sums. The same run in
FP64 fields is timed too, and the FP32 k_s error against it is reported.
``--workload c4`` runs the 5 um full electrode (1066 x 1066 x 410,
4580 steps, the reference's published A100 headline).

Prints one JSON line. ``value`` is interior voxel-steps per second over the
timed region; ``roofline`` divides the sweep kernel's compulsory HBM bytes
(read n+ and n-, write n+' and n-': 16 B per interior voxel in FP32, 32 B in
FP64) by its average duration, measured with HIP events around every sweep
launch in the timed region; ``cpu_baseline`` times the CPU oracle (a C/OpenMP
restatement of the reference's batched backend) on a bounded prefix of the
same run.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT / "iontracks-pulsedproton-python_amd"), str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "voxel-steps/s & % HBM BW; batched k_s-curve time vs plain at |dk_s|<=1e-6"
PEAK_GBS = 8000.0


def workload(name):
    from configs import aic_config, readme_config

    if name == "c1":
        return readme_config(5.0, 0.05), "5um_r0.05cm_mid_solve"
    if name == "c0":
        return readme_config(10.0, 0.008), "10um_r0.008cm_readme_example"
    if name == "c4":
        return aic_config(0.265, 5.0), "5um_full_electrode_r0.265cm"
    raise SystemExit(f"unknown workload {name!r} (c0, c1, c4)")


def interior_voxels(cfg):
    return (cfg.no_xy - 2) ** 2 * (cfg.no_z_with_buffer - 2)


def gpu_run(cfg, schedule, xs, ys, precision, steps, warmup):
    from pulsed_ion_chamber_hip.solver_hip import HipSimulation
    from pulsed_ion_chamber_hip.state import collection_efficiency

    with HipSimulation(cfg, precision) as sim:
        sim.upload_tracks(schedule, xs, ys)
        if warmup:
            sim.run_steps(0, min(warmup, cfg.total_time_steps))
            sim.synchronize()
        sim.reset()
        sim.enable_timing(True)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run_steps(0, steps)
        sim.synchronize()
        elapsed = time.perf_counter() - t0
        timing = sim.timing()
        per_step = sim.read_per_step(0, steps)
    f_t = collection_efficiency(per_step[:, 0], per_step[:, 1], steps)
    ks = 1.0 / f_t[-1] if steps == cfg.total_time_steps else float("nan")
    return elapsed, timing, ks


def traffic_from_profiles(kernel_tag):
    """HBM bytes per sweep launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None if not profiled."""
    path = ROOT / "profiles" / f"{kernel_tag}_traffic.json"
    if not path.exists():
        return None
    try:
        return float(json.loads(path.read_text())["hbm_bytes_per_launch"])
    except (ValueError, KeyError):
        return None


def cpu_baseline(cfg, schedule, xs, ys, budget_s):
    from oracle import batched

    threads = min(16, len(os.sched_getaffinity(0)))
    nvox = interior_voxels(cfg)
    probe = min(20, cfg.total_time_steps)
    t0 = time.perf_counter()
    batched.run_per_step(cfg, schedule, xs, ys, n_steps=probe, nthreads=threads, fields=False)
    rate = probe / (time.perf_counter() - t0)
    sample = int(min(cfg.total_time_steps, max(probe, rate * budget_s)))
    t0 = time.perf_counter()
    batched.run_per_step(cfg, schedule, xs, ys, n_steps=sample, nthreads=threads, fields=False)
    elapsed = time.perf_counter() - t0
    return {
        "value": nvox * sample / elapsed,
        "unit": "voxel-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample} of {cfg.total_time_steps} steps of the same run (deposition + sweep + ring), "
                  f"FP64 fields, {threads} OpenMP threads, {elapsed:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the whole run)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c4"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-reference-precision", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from pulsed_ion_chamber_hip import _lib
    from pulsed_ion_chamber_hip.pulses import draw_tracks

    device, _, _ = _lib.device_info()
    cfg, wl_name = workload(args.workload)
    total = cfg.total_time_steps
    steps = total if args.steps <= 0 else min(args.steps, total)
    t_host = time.perf_counter()
    schedule, xs, ys = draw_tracks(cfg)
    host_draw_s = time.perf_counter() - t_host
    nvox = interior_voxels(cfg)
    print(f"bench: {wl_name} {cfg.no_xy}x{cfg.no_xy}x{cfg.no_z_with_buffer}, {steps}/{total} steps, "
          f"{len(xs)} tracks, {args.precision}, device {device}", file=sys.stderr, flush=True)

    elapsed, timing, ks = gpu_run(cfg, schedule, xs, ys, args.precision, steps, args.warmup)
    bytes_per_voxel = 16 if args.precision == "fp32" else 32
    sweep_avg_ms = timing["sweep_ms"] / max(1, timing["sweep_launches"])
    achieved = nvox * bytes_per_voxel / (sweep_avg_ms * 1e-3) / 1e9
    tag = f"sweep_{args.workload}_{args.precision}"
    traffic = traffic_from_profiles(tag)
    out = {
        "metric": METRIC,
        "value": nvox * steps / elapsed,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": wl_name,
            "grid": [cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer],
            "interior_voxels": nvox,
            "total_time_steps": total,
            "tracks": int(len(xs)),
            "fields": args.precision,
            "accumulators": "f64 (recombination product, deposition, scored sums)",
            "seed": cfg.seed,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_GBS,
            "traffic": traffic,
            "kernel": "sweep_kernel",
            "bytes_per_interior_voxel": bytes_per_voxel,
            "avg_launch_ms": sweep_avg_ms,
            "launches": timing["sweep_launches"],
        },
        "breakdown_ms_per_step": {
            "sweep": timing["sweep_ms"] / steps,
            "deposit": timing["deposit_ms"] / steps,
            "ring_and_reduce": timing["ring_reduce_ms"] / steps,
        },
        "ks": ks,
        "host_track_draw_s": host_draw_s,
    }
    if not args.no_reference_precision and args.precision == "fp32":
        e64, t64, ks64 = gpu_run(cfg, schedule, xs, ys, "fp64", steps, 0)
        sw64 = t64["sweep_ms"] / max(1, t64["sweep_launches"])
        out["fp64"] = {
            "value": nvox * steps / e64,
            "ms_per_step": e64 / steps * 1e3,
            "ks": ks64,
            "sweep_avg_ms": sw64,
            "sweep_GBs": nvox * 32 / (sw64 * 1e-3) / 1e9,
        }
        out["dks_fp32_vs_fp64"] = abs(ks - ks64) if steps == total else None
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, schedule, xs, ys, args.cpu_budget_s)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
