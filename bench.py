"""Benchmark of the pulsed-ion-chamber time-stepping loop on one MI355X.

One "step" is one simulation time step over the whole grid: the gather
deposition of that step's tracks (in-pulse steps), the fused Lax-Wendroff
sweep of both species with its scored reduction, the lateral-ring update and
the per-step reduction. The timed region starts with every track of the run
already resident in HBM and ends when the last step's record is written.

Default workload (the headline): the 5 um full electrode of the Classic
Markus chamber -- AIC-144 physics, 56.2 MeV/u protons, 300 V over 0.2 cm,
540 us pulses at 50 Hz, 10 Gy/s to water, r 0.265 cm, 1066 x 1066 x 410 =
466 M voxels, 4580 steps, 24.6 M tracks, seed 20260527 -- with FP32 fields
and FP64 recombination and scored sums. ``--steps K`` times K steps from the
start of the run; outside the timed region the whole run is then done in
FP32 and in FP64 to report k_s, the FP32 k_s error and the FP64 k_s against
the reference's published value (1.1104725894530176 on its GPU backend).

Other workloads: ``c1`` the mid-size solve (library-default physics, 5 um
over r 0.05 cm, 208 x 208 x 406, 3388 steps, seed 1; also reported as the
``mid_size`` key of the default line); ``c0`` the README example (24 x 24 x
206); ``c5`` the 2 um full electrode (2656 x 2656 x 1010 = 7.1 G voxels,
114 GB of FP32 fields; give it --steps, the whole run is 13 019 steps);
``c2`` the 32-point saturation curve (8 voltages 50-500 V x 4 dose rates,
grids 16^2..47^2 x 106..206, 243..2505 steps) timed as one batch against 32
sequential plain runs of the same configurations (the ``curve`` key of the
default line); ``c3`` the adjoint (dk_s/dtheta over 3 pulses of a 48^2 x 206
grid, FP64) with its Richardson-extrapolated finite-difference check;
``published`` every run the reference publishes a wall time for, rerun whole
in FP64 and FP32 beside that time and its published k_s or f.

Prints one JSON line. ``value`` is interior voxel-steps per second over the
timed region; ``roofline`` divides the sweep kernel's compulsory HBM bytes
(read n+ and n-, write n+' and n-': 16 B per interior voxel and step in FP32,
32 B in FP64; a launch of the two-step pass, which FP32 grids with the
reflecting wall take, counts both steps) by its average duration, measured with HIP events on the handle's
stream over the timed region: one event pair around each run of 16
back-to-back sweep launches, so every launch is timed (with the ~1 us
between launches) while the pairs cost the run under 1 %; ``traffic`` is
the copy-calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE of the same kernel from
the committed profile. ``cpu_baseline`` times the CPU oracle (a C/OpenMP
restatement of the reference's batched backend, compiled here with
-march=native) on a bounded prefix of the same run at 1, 2, 4, ... threads up
to the cores this process may use, and reports the best, with the host's CPU
model, NUMA node count and L3 size.
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


def workload(name, dose_rate=None):
    from configs import aic_config, readme_config

    if name == "c1":
        return readme_config(5.0, 0.05), "5um_r0.05cm_mid_solve"
    if name == "c0":
        return readme_config(10.0, 0.008), "10um_r0.008cm_readme_example"
    if name == "c4":
        if dose_rate is not None:
            return (aic_config(0.265, 5.0, dose_rate_water_Gy_s=dose_rate),
                    f"5um_full_electrode_r0.265cm_{dose_rate:g}Gy_s")
        return aic_config(0.265, 5.0), "5um_full_electrode_r0.265cm"
    if name == "c5":
        return aic_config(0.265, 2.0, max_voxels=1e10), "2um_full_electrode_r0.265cm"
    raise SystemExit(f"unknown workload {name!r} (c0, c1, c4, c5)")


# sweep launches per pair of timing events
TIMING_GROUP = 16


def interior_voxels(cfg):
    return (cfg.no_xy - 2) ** 2 * (cfg.no_z_with_buffer - 2)


def interior_voxels_of(run):
    nx, _, nz = run["grid"]
    return (nx - 2) ** 2 * (nz - 2)


def gpu_run(cfg, schedule, xs, ys, precision, steps, warmup):
    from pulsed_ion_chamber_hip.solver_hip import HipSimulation
    from pulsed_ion_chamber_hip.state import collection_efficiency

    with HipSimulation(cfg, precision) as sim:
        sim.upload_tracks(schedule, xs, ys)
        if warmup:
            sim.run_steps(0, min(warmup, cfg.total_time_steps))
            sim.synchronize()
        sim.reset()
        sim.enable_timing(True, group=TIMING_GROUP)
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


def host_description():
    """The CPU the baseline runs on: logical CPUs, the ones this process may
    use, the thread cap the environment sets, and lscpu's model, sockets,
    NUMA nodes and L3."""
    import subprocess

    usable = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(usable, int(cap)) if cap.isdigit() and int(cap) > 0 else usable
    info = {"nproc": os.cpu_count(), "usable_cpus": usable, "omp_num_threads_env": cap or None,
            "threads_available": threads}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        text = ""
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "NUMA node(s)": "numa_nodes", "L3 cache": "l3"}
    for line in text.splitlines():
        name, _, value = line.partition(":")
        if name.strip() in keys and keys[name.strip()] not in info:
            info[keys[name.strip()]] = value.strip()
    return info


def cpu_baseline(cfg, schedule, xs, ys, budget_s):
    """The CPU oracle on the first steps of the same run, at 1, 2, 4, ...
    threads up to every core this process may use (the reference shows that
    fewer threads can win on some grids, docs/PERFORMANCE.md:180-202); the
    best rate is the baseline. Each count runs a step budget sized from a
    pilot at full width, so the leg takes about ``budget_s`` of loop time;
    the loop is timed inside the oracle, without its allocation and first
    touch. The oracle is compiled for this host (-march=native) when gcc
    can, else the in-tree x86-64-v3 build is used."""
    import subprocess
    import tempfile

    from oracle import batched

    host = host_description()
    try:
        batched.use_library(batched.build_native(Path(tempfile.mkdtemp(prefix="pic_oracle_native_"))))
        build = "-O3 -march=native"
    except (OSError, subprocess.SubprocessError):
        build = "-O3 -march=x86-64-v3 (in-tree build; native compile failed)"
    nvox = interior_voxels(cfg)
    total = cfg.total_time_steps
    top = host["threads_available"]
    counts = sorted({1 << k for k in range(top.bit_length()) if (1 << k) <= top} | {top})
    pilot = min(2, total)
    batched.run_per_step(cfg, schedule, xs, ys, n_steps=pilot, nthreads=top, fields=False)
    rate_top = pilot / max(batched.last_loop_seconds(), 1e-9)
    share = budget_s / len(counts)
    sweep = []
    for t in counts:
        n = int(min(total, max(1, share * rate_top * t / top)))
        batched.run_per_step(cfg, schedule, xs, ys, n_steps=n, nthreads=t, fields=False)
        sec = batched.last_loop_seconds()
        sweep.append({"threads": t, "steps": n, "loop_s": sec, "voxel_steps_per_s": nvox * n / sec})
    best = max(sweep, key=lambda r: r["voxel_steps_per_s"])
    return {
        "value": best["voxel_steps_per_s"],
        "unit": "voxel-steps/s",
        "cores": best["threads"],
        "kind": "port",
        "sample": f"first {best['steps']} of {total} steps of the same run (deposition + sweep + ring), FP64 fields, "
                  f"{best['threads']} OpenMP threads (best of {counts}), {best['loop_s']:.1f} s of loop",
        "oracle_build": build,
        "threads_sweep": sweep,
        "host": host,
    }


def curve_bench(precision, cpu_budget_s):
    """The 32-point saturation curve: one batch against 32 sequential plain
    runs of the same configurations, tracks already resident in HBM for both
    (the host draws them once, untimed), each side the median of three timed
    passes. Returns the curve's figures."""
    from configs import saturation_curve_configs
    from pulsed_ion_chamber_hip.batch import HipBatch
    from pulsed_ion_chamber_hip.pulses import draw_tracks
    from pulsed_ion_chamber_hip.solver_hip import HipSimulation
    from pulsed_ion_chamber_hip.state import collection_efficiency

    cfgs = saturation_curve_configs()
    draws = [draw_tracks(c) for c in cfgs]
    vsteps = [interior_voxels(c) * c.total_time_steps for c in cfgs]

    def ks_of(sim, c):
        per = sim.read_per_step(0, c.total_time_steps)
        return 1.0 / collection_efficiency(per[:, 0], per[:, 1], c.total_time_steps)[-1]

    # Both sides are timed in steady state: one whole untimed pass first, so
    # every kernel and tiling either side uses has been launched once (a
    # kernel's first launch loads its code) and every buffer is allocated;
    # then PASSES timed passes of each, whose medians are compared (one pass
    # of the batch varies by about a millisecond from the next).
    PASSES = 3

    def batched(prec, timed, passes):
        sims = [HipSimulation(c, prec) for c in cfgs]
        for sim, d in zip(sims, draws):
            sim.upload_tracks(*d)
        loops = []
        with HipBatch(sims) as b:
            b.run()
            b.synchronize()
            for _ in range(passes):
                for sim in sims:
                    sim.reset()
                b.synchronize()
                t0 = time.perf_counter()
                b.run()
                b.synchronize()
                loops.append(time.perf_counter() - t0)
            timing = None
            if timed:
                # kernel times from one more pass with events around every
                # launch (the events cost host time, so not in the wall clock)
                for sim in sims:
                    sim.reset()
                b.enable_timing(True)
                b.run()
                timing = b.timing()
            tiling = b.tiling()
        ks = [ks_of(sim, c) for sim, c in zip(sims, cfgs)]
        for sim in sims:
            sim.close()
        return loops, ks, timing, tiling

    def sequential(prec, passes):
        sims = [HipSimulation(c, prec) for c in cfgs]
        try:
            for sim, c, d in zip(sims, cfgs, draws):  # the untimed pass
                sim.upload_tracks(*d)
                sim.run_steps(0, c.total_time_steps)
                sim.synchronize()
            totals = []
            for _ in range(passes):
                total = 0.0
                for sim, c in zip(sims, cfgs):
                    sim.reset()
                    sim.synchronize()
                    t0 = time.perf_counter()
                    for first in range(0, c.total_time_steps, 256):  # run_simulation_hip's chunking
                        sim.run_steps(first, min(256, c.total_time_steps - first))
                    sim.synchronize()
                    total += time.perf_counter() - t0
                totals.append(total)
            ks = [ks_of(sim, c) for sim, c in zip(sims, cfgs)]
        finally:
            for sim in sims:
                sim.close()
        return totals, ks

    loops_b, ks_b, timing, tiling = batched(precision, True, PASSES)
    loops_s, ks_s = sequential(precision, PASSES)
    loop_b, loop_s = float(np.median(loops_b)), float(np.median(loops_s))
    _, ks64, _, _ = batched("fp64", False, 0)
    bytes_per_voxel = 16 if precision == "fp32" else 32
    sweep_bytes = sum(vsteps) * bytes_per_voxel
    out = {
        "configs": len(cfgs),
        "voxel_steps": sum(vsteps),
        "global_steps": max(c.total_time_steps for c in cfgs),
        "grids": f"{min(c.no_xy for c in cfgs)}^2..{max(c.no_xy for c in cfgs)}^2 x "
                 f"{min(c.no_z_with_buffer for c in cfgs)}..{max(c.no_z_with_buffer for c in cfgs)}",
        "precision": precision,
        "batched_s": loop_b,
        "sequential_plain_s": loop_s,
        "speedup": loop_s / loop_b,
        "timed_passes": PASSES,
        "batched_s_passes": loops_b,
        "sequential_plain_s_passes": loops_s,
        "voxel_steps_per_s": sum(vsteps) / loop_b,
        "batched_sweep_ms": timing["sweep_ms"],
        "batched_sweep_launches": timing["sweep_launches"],
        "batched_sweep_GBs": sweep_bytes / (timing["sweep_ms"] * 1e-3) / 1e9,
        "gather_ms": timing["gather_ms"],
        "gather_launches": timing["gather_launches"],
        "tiling": tiling,
        "max_dks_vs_fp64": max(abs(a - b) for a, b in zip(ks_b, ks64)),
        "max_dks_batched_vs_plain": max(abs(a - b) for a, b in zip(ks_b, ks_s)),
    }
    if cpu_budget_s > 0:
        from oracle import batched as oracle_batched

        threads = host_description()["threads_available"]
        t0 = time.perf_counter()
        done = 0
        dks = rel64 = 0.0
        for c, d, k64 in zip(cfgs, draws, ks64):
            per, _, _ = oracle_batched.run_per_step(c, *d, nthreads=threads, fields=False)
            k_ref = 1.0 / collection_efficiency(per[:, 0], per[:, 1], c.total_time_steps)[-1]
            dks = max(dks, abs(ks_b[done] - k_ref))
            rel64 = max(rel64, abs(k64 - k_ref) / k_ref)
            done += 1
            if time.perf_counter() - t0 > cpu_budget_s:
                break
        elapsed = time.perf_counter() - t0
        out["max_dks_vs_oracle"] = dks
        out["max_rel_dks_fp64_vs_oracle"] = rel64
        out["cpu_oracle"] = {
            "configs": done,
            "seconds": elapsed,
            "voxel_steps_per_s": sum(vsteps[:done]) / elapsed,
            "threads": threads,
        }
    return out


def published_runs_bench(precisions=("fp64", "fp32")):
    """Every run the reference publishes a wall time for (the 10 um ladder
    and the 5 um full electrode on an A100, r 0.05 cm at 10, 5 and 2 um on a
    GH200, and the GH200's 200-step resolution ladder and 500 steps of the
    2 um full electrode), rerun here with run_simulation_hip. Like the
    reference's wall_s, each time is the whole call -- host track draw,
    allocation, upload, the steps and the read-back -- after one warm-up
    call that loads the kernels."""
    from configs import aic_config
    from pulsed_ion_chamber_hip import run_simulation_hip

    table = json.loads((ROOT / "tests" / "golden" / "published.json").read_text())
    cases, seen = [], set()
    for row in table["aic_runs"]:
        key = (row["sampled_radius_cm"], row["grid_size_um"])
        if "gpu_wall_s" in row and key not in seen:
            seen.add(key)
            cases.append((row, None))
    for row in table["prefix_f"]:
        if "gpu_wall_s" in row:
            cases.append((row, row["steps"]))
    run_simulation_hip(aic_config(0.014, 10.0), progress=False, return_fields=False)  # warm-up
    out = []
    for row, max_steps in cases:
        cfg = aic_config(row["sampled_radius_cm"], row["grid_size_um"], max_voxels=1e10)
        steps = max_steps or cfg.total_time_steps
        rec = {"sampled_radius_cm": row["sampled_radius_cm"], "grid_size_um": row["grid_size_um"],
               "grid": [cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer], "steps": steps,
               "reference_wall_s": row["gpu_wall_s"], "reference_gpu": row["gpu"]}
        if "cpu_wall_s" in row:
            rec["reference_cpu_wall_s"] = row["cpu_wall_s"]
        for prec in precisions:
            print(f"bench: published run r {row['sampled_radius_cm']} cm @ {row['grid_size_um']} um, {steps} steps, "
                  f"{prec}", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            res = run_simulation_hip(cfg, progress=False, precision=prec, return_fields=False, max_steps=max_steps)
            wall = time.perf_counter() - t0
            rec[f"{prec}_wall_s"] = wall
            rec[f"{prec}_speedup_vs_reference"] = row["gpu_wall_s"] / wall
            if max_steps:
                f = float(res.f_t[steps - 1])
                rec[f"{prec}_f_end"] = f
                rec[f"{prec}_abs_err_f_end"] = abs(f - row["f_end"])
            else:
                rec[f"{prec}_ks"] = float(res.ks)
                rec[f"{prec}_abs_err_ks"] = abs(float(res.ks) - row["gpu_ks"])
        out.append(rec)
    return out


def adjoint_bench(fd: bool = True, pulses: int = 3):
    """dk_s/dtheta on the 10 um, r 0.02 cm grid over 3 pulses (FP64 fields,
    segments of 64 steps, their starts kept in two levels): forward and
    reverse time (the reverse includes one more forward pass that recomputes
    the segment starts), device memory, and
    the gradient against Richardson-extrapolated FP64 central differences
    (relative steps 1e-3 and 2e-3, dt and step count frozen) on all six
    parameters."""
    from configs import readme_config
    from pulsed_ion_chamber_hip.adjoint import PARAMS, KsModel

    cfg = readme_config(10.0, 0.02, n_pulses=pulses)
    model = KsModel(cfg, checkpoint_every=64)
    model.ks_and_gradient()  # warm-up (kernels loaded, memory touched)
    t0 = time.perf_counter()
    ks, grad = model.ks_and_gradient()
    wall = time.perf_counter() - t0
    info = dict(model.last_info)
    t0 = time.perf_counter()
    ks_plain = model.ks(model.theta0)
    plain = time.perf_counter() - t0
    out = {
        "grid": [cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer],
        "steps": cfg.total_time_steps,
        "ks": ks,
        "ks_forward_only": ks_plain,
        "dks_dtheta": dict(zip(PARAMS, grad.tolist())),
        "theta": dict(zip(PARAMS, model.theta0.tolist())),
        "forward_with_checkpoints_s": info["forward_ms"] / 1e3,
        "reverse_s": info["backward_ms"] / 1e3,
        "reverse_over_forward": info["backward_ms"] / info["forward_ms"],
        "gradient_wall_s": wall,
        "plain_forward_wall_s": plain,
        "checkpoint_every": info["checkpoint_every"],
        "segments": info["segments"],
        "segments_per_outer_checkpoint": info["segments_per_outer"],
        "outer_checkpoints": info["outer_checkpoints"],
        "adjoint_device_bytes": info["device_bytes"],
        "plain_state_bytes": 4 * cfg.no_xy**2 * cfg.no_z_with_buffer * 8,
    }
    out["memory_over_plain"] = out["adjoint_device_bytes"] / out["plain_state_bytes"]
    if fd:
        t0 = time.perf_counter()
        fdv = model.finite_difference_richardson(rel=1e-3)
        out["fd_s"] = time.perf_counter() - t0
        out["fd_method"] = "Richardson-extrapolated central differences, (4 D(h) - D(2h)) / 3, h = 1e-3 theta"
        out["fd_dks_dtheta"] = dict(zip(PARAMS, fdv.tolist()))
        rel = np.abs(grad - fdv) / np.abs(fdv)
        out["rel_err_vs_fd"] = dict(zip(PARAMS, rel.tolist()))
        out["max_rel_err_vs_fd"] = float(np.max(rel))
    return cfg, out



def gather_batch_steps(cfg):
    """Steps per gather launch, as the library sizes it: as many as fit in
    256 MB of FP64 density, at most 128, even where more than one fits."""
    n = max(1, min(128, (256 << 20) // (cfg.no_xy * cfg.no_xy * 8)))
    return n & ~1 if n > 1 else n


def deposition_alone(cfg, schedule, xs, ys, precision, steps):
    """The first ``steps`` steps with graph replay: each segment's bin passes
    and gather are then launched in the sweeps' own stream, so under
    rocprofv3 --kernel-trace their durations are the deposition's with the
    device to itself (in a timed run they share it with the sweeps)."""
    from pulsed_ion_chamber_hip.solver_hip import HipSimulation

    with HipSimulation(cfg, precision) as sim:
        sim.set_graphs(True)
        sim.upload_tracks(schedule, xs, ys)
        t0 = time.perf_counter()
        sim.run_steps(0, steps)
        sim.synchronize()
        wall = time.perf_counter() - t0
        per = sim.read_per_step(0, steps)
    return {"steps": steps, "wall_s": wall, "injected": float(per[:, 0].sum()),
            "gather_batch_steps": gather_batch_steps(cfg)}


def single_run_line(cfg, workload_key, schedule, xs, ys, precision, steps, warmup):
    """Time ``steps`` steps from the start of the run (after ``warmup``
    untimed steps and a reset) with every track resident in HBM; the sweep's
    roofline from its HIP-event timing over the same region."""
    nvox = interior_voxels(cfg)
    elapsed, timing, ks = gpu_run(cfg, schedule, xs, ys, precision, steps, warmup)
    bytes_per_voxel = 16 if precision == "fp32" else 32
    launches = max(1, timing["sweep_launches"])
    sweep_avg_ms = timing["sweep_ms"] / launches
    sweep_step_ms = timing["sweep_ms"] / steps
    # a launch of the two-step pass (FP32, reflecting wall) advances two
    # steps: its algorithmic bytes are those of both, 2 x 16 B per voxel
    steps_per_launch = steps / launches
    achieved = nvox * bytes_per_voxel * steps_per_launch / (sweep_avg_ms * 1e-3) / 1e9
    return {
        "value": nvox * steps / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "wall_s": elapsed,
        "ks": ks,
        "timing": timing,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_GBS,
            "traffic": traffic_from_profiles(f"sweep_{workload_key}_{precision}"),
            "kernel": "sweep2_kernel" if steps_per_launch > 1.5 else "sweep_kernel",
            "bytes_per_interior_voxel_step": bytes_per_voxel,
            "steps_per_launch": steps_per_launch,
            "avg_launch_ms": sweep_avg_ms,
            "sweep_ms_per_step": sweep_step_ms,
            "launches": timing["sweep_launches"],
            "launches_per_event_pair": TIMING_GROUP,
        },
        "breakdown_ms_per_step": {
            # a segment's gather (one launch per batch of steps) runs on a
            # second stream beside the previous segment's sweeps; with a
            # single segment there is nothing beside which to run
            "deposit_overlaps_sweep": steps > gather_batch_steps(cfg),
            "sweep": sweep_step_ms,
            "deposit": timing["deposit_ms"] / steps,
            "deposit_launches": timing["deposit_launches"],
            "ring_and_reduce": timing["ring_reduce_ms"] / steps,
        },
    }


def published_ks(workload_key):
    """The reference's published full-precision k_s for a workload, from the
    committed fixture of its data files (profiling/data/gpu_athena/*.json)."""
    where = {"c4": (0.265, 5.0)}.get(workload_key)
    if where is None:
        return None
    table = json.loads((ROOT / "tests" / "golden" / "published.json").read_text())
    for row in table["aic_runs"]:
        if (row["sampled_radius_cm"], row["grid_size_um"]) == where:
            return {"gpu_ks": row["gpu_ks"], "cpu_ks": row.get("cpu_ks"),
                    "source": "profiling/data/gpu_athena/gpu_5um_full_electrode.json (A100, FP64, 156.3 s)"}
    return None


def whole_runs(cfg, workload_key, schedule, xs, ys, precision, timed_line, with_fp64):
    """The whole run in the benchmark precision (reusing the timed one when
    it was whole) and, for FP32, in FP64: k_s of each, the FP32 error and the
    FP64 value against the reference's published one. Untimed for ``value``;
    each run's wall time is reported."""
    total = cfg.total_time_steps
    nvox = interior_voxels(cfg)
    run = timed_line or single_run_line(cfg, workload_key, schedule, xs, ys, precision, total, 0)
    out = {"ks": run["ks"], "full_run": {
        "precision": precision, "steps": total, "wall_s": run["wall_s"], "ms_per_step": run["ms_per_step"],
        "sweep_ms_per_step": run["roofline"]["sweep_ms_per_step"], "sweep_frac_of_peak": run["roofline"]["frac"],
        "ks": run["ks"]}}
    ks64 = run["ks"] if precision == "fp64" else None
    if with_fp64 and precision == "fp32":
        e64, t64, ks64 = gpu_run(cfg, schedule, xs, ys, "fp64", total, 0)
        sw64 = t64["sweep_ms"] / total
        out["fp64"] = {"value": nvox * total / e64, "wall_s": e64, "ms_per_step": e64 / total * 1e3, "ks": ks64,
                       "sweep_ms_per_step": sw64, "sweep_GBs": nvox * 32 / (sw64 * 1e-3) / 1e9}
        out["dks_fp32_vs_fp64"] = abs(run["ks"] - ks64)
    pub = published_ks(workload_key)
    if pub is not None:
        pub = dict(pub)
        if ks64 is not None:
            pub["rel_err_fp64_vs_gpu_ks"] = abs(ks64 - pub["gpu_ks"]) / pub["gpu_ks"]
        if precision == "fp32":
            pub["abs_err_fp32_vs_gpu_ks"] = abs(run["ks"] - pub["gpu_ks"])
        out["ks_published"] = pub
    return out


def mid_size_line():
    """The mid-size solve (5 um, r 0.05 cm, 208^2 x 406, 3388 steps), whole
    run in FP32 (timed) and FP64 (for the k_s error)."""
    from pulsed_ion_chamber_hip.pulses import draw_tracks

    cfg, name = workload("c1")
    schedule, xs, ys = draw_tracks(cfg)
    line = single_run_line(cfg, "c1", schedule, xs, ys, "fp32", cfg.total_time_steps, 50)
    e64, t64, ks64 = gpu_run(cfg, schedule, xs, ys, "fp64", cfg.total_time_steps, 0)
    return {"workload": name, "grid": [cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer],
            "steps": cfg.total_time_steps, "value": line["value"], "ms_per_step": line["ms_per_step"],
            "roofline": line["roofline"], "ks": line["ks"], "ks_fp64": ks64, "dks_fp32_vs_fp64": abs(line["ks"] - ks64)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the whole run)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c4", choices=["c0", "c1", "c2", "c3", "c4", "c5", "published"])
    ap.add_argument("--no-curve", action="store_true", help="skip the saturation-curve comparison")
    ap.add_argument("--adjoint-pulses", type=int, default=3, help="pulses of the adjoint workload (c3)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-reference-precision", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-full-run", action="store_true", help="skip the whole runs that give k_s")
    ap.add_argument("--no-mid-size", action="store_true", help="skip the mid-size solve key of the c4 line")
    ap.add_argument("--cpu-budget-s", type=float, default=25.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dose-rate", type=float, default=None,
                    help="water dose rate of the c4 workload in Gy/s for --deposition-alone (default: its 10 Gy/s)")
    ap.add_argument("--deposition-alone", action="store_true",
                    help="only run --steps steps (default 112) with graph replay, the deposition alone on the device "
                         "(for rocprofv3 kernel statistics), and print their wall time")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from pulsed_ion_chamber_hip import _lib
    from pulsed_ion_chamber_hip.pulses import draw_tracks

    device, _, _ = _lib.device_info()
    if args.workload == "c2":
        print(f"bench: 32-point saturation curve, {args.precision}, device {device}", file=sys.stderr, flush=True)
        curve = curve_bench(args.precision, 0.0 if args.no_cpu else args.cpu_budget_s)
        out = {
            "metric": METRIC,
            "value": curve["voxel_steps_per_s"],
            "unit": "voxel-steps/s",
            "n_gpus": 1,
            "steps": curve["global_steps"],
            "warmup": curve["global_steps"],  # one whole untimed pass
            "ms_per_step": curve["batched_s"] / curve["global_steps"] * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32" if args.precision == "fp32" else "f64",
            "data": "synthetic",
            "config": {"workload": "saturation_curve_32", "grids": curve["grids"], "fields": args.precision},
            "roofline": {
                "bound": "hbm",
                "achieved": curve["batched_sweep_GBs"],
                "peak": PEAK_GBS,
                "unit": "GB/s",
                "frac": curve["batched_sweep_GBs"] / PEAK_GBS,
                "traffic": None,
                "kernel": "batch_sweep_kernel",
                "avg_launch_ms": curve["batched_sweep_ms"] / max(1, curve["batched_sweep_launches"]),
                "launches": curve["batched_sweep_launches"],
            },
            "curve": curve,
        }
        if "cpu_oracle" in curve:
            co = curve["cpu_oracle"]
            out["cpu_baseline"] = {
                "value": co["voxel_steps_per_s"], "unit": "voxel-steps/s", "cores": co["threads"], "kind": "port",
                "sample": f"first {co['configs']} of 32 curve configurations, whole runs, FP64 fields, "
                          f"{co['threads']} OpenMP threads, {co['seconds']:.1f} s",
            }
        print(json.dumps(out))
        return
    if args.workload == "published":
        print(f"bench: the reference's published runs, device {device}", file=sys.stderr, flush=True)
        runs = published_runs_bench()
        vs = sum(interior_voxels_of(r) * r["steps"] for r in runs)
        wall64 = sum(r["fp64_wall_s"] for r in runs)
        out = {
            "metric": METRIC,
            "value": vs / wall64,
            "unit": "voxel-steps/s",
            "n_gpus": 1,
            "steps": sum(r["steps"] for r in runs),
            "warmup": 1,
            "ms_per_step": wall64 / sum(r["steps"] for r in runs) * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "published_runs", "runs": len(runs),
                       "timing": "whole run_simulation_hip call, as the reference's wall_s"},
            "runs": runs,
        }
        print(json.dumps(out))
        return
    if args.workload == "c3":
        print(f"bench: adjoint dk_s/dtheta, 10 um r 0.02 cm, 3 pulses, fp64, device {device}", file=sys.stderr,
              flush=True)
        cfg, adj = adjoint_bench(fd=not args.no_cpu, pulses=args.adjoint_pulses)
        nvox = interior_voxels(cfg)
        total_s = adj["forward_with_checkpoints_s"] + adj["reverse_s"]
        out = {
            "metric": METRIC,
            "value": nvox * cfg.total_time_steps / adj["forward_with_checkpoints_s"],
            "unit": "voxel-steps/s",
            "n_gpus": 1,
            "steps": cfg.total_time_steps,
            "warmup": 1,
            "ms_per_step": total_s / cfg.total_time_steps * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"adjoint_10um_r0.02cm_{args.adjoint_pulses}pulses", "grid": adj["grid"], "fields": "fp64",
                       "checkpoint_every": adj["checkpoint_every"],
                       "segments_per_outer_checkpoint": adj["segments_per_outer_checkpoint"]},
            "adjoint": adj,
        }
        print(json.dumps(out))
        return
    if args.dose_rate is not None and not (args.deposition_alone and args.workload == "c4"):
        ap.error("--dose-rate applies to the c4 workload with --deposition-alone")
    cfg, wl_name = workload(args.workload, args.dose_rate)
    total = cfg.total_time_steps
    steps = total if args.steps <= 0 else min(args.steps, total)
    if args.deposition_alone:
        n = min(112, total) if args.steps <= 0 else steps
        schedule, xs, ys = draw_tracks(cfg)
        out = {"workload": wl_name, "deposition_alone": deposition_alone(cfg, schedule, xs, ys, args.precision, n)}
        print(json.dumps(out))
        return
    if args.workload == "c5":
        # the FP64 comparison run would need 228 GB of device memory, the CPU
        # oracle 228 GB of host memory, and the whole FP32 run ~6 minutes:
        # only the timed steps run
        args.no_cpu = True
        args.no_reference_precision = True
        args.no_full_run = steps < total
    t_host = time.perf_counter()
    schedule, xs, ys = draw_tracks(cfg)
    host_draw_s = time.perf_counter() - t_host
    print(f"bench: {wl_name} {cfg.no_xy}x{cfg.no_xy}x{cfg.no_z_with_buffer}, {steps}/{total} steps, "
          f"{len(xs)} tracks, {args.precision}, device {device}", file=sys.stderr, flush=True)

    line = single_run_line(cfg, args.workload, schedule, xs, ys, args.precision, steps, args.warmup)
    out = {
        "metric": METRIC,
        "value": line["value"],
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": line["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": wl_name,
            "grid": [cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer],
            "interior_voxels": interior_voxels(cfg),
            "total_time_steps": total,
            "tracks": int(len(xs)),
            "fields": args.precision,
            "accumulators": "f64 (recombination product, deposition, scored sums)",
            "seed": cfg.seed,
        },
        "roofline": line["roofline"],
        "breakdown_ms_per_step": line["breakdown_ms_per_step"],
        "ks": line["ks"],
        "host_track_draw_s": host_draw_s,
    }
    if not args.no_full_run:
        print("bench: whole runs for k_s", file=sys.stderr, flush=True)
        out.update(whole_runs(cfg, args.workload, schedule, xs, ys, args.precision,
                              line if steps == total else None, not args.no_reference_precision))
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, schedule, xs, ys, args.cpu_budget_s)
    if args.workload == "c4" and not args.no_mid_size:
        print("bench: mid-size solve", file=sys.stderr, flush=True)
        out["mid_size"] = mid_size_line()
    if not args.no_curve and args.workload in ("c1", "c4"):
        print("bench: saturation curve", file=sys.stderr, flush=True)
        out["curve"] = curve_bench("fp32", 0.0 if args.no_cpu else args.cpu_budget_s)
    print(json.dumps(out))

if __name__ == "__main__":
    main()
