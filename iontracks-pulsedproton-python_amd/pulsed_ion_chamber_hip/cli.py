"""IFJ PAN AIC-144, PTW Markus 23343 (2 mm gap), macropulse, on the GPU.

The reference's tier driver (examples/ifj_aic144/run_markus_2mm.py) with the
HIP backend: same scenario, tiers, options and outputs, so a record written
here plots with the reference's examples/ifj_aic144/plot.py.

    python -m pulsed_ion_chamber_hip.cli [tier] [--precision fp64|fp32]
        [--dose-rate-water-Gy-s R] [--sampled-radius-mm R] [--json FILE]
        [--save-run DIR] [--dry-run | --estimate-runtime-seconds N]
        [--tiling auto]
"""

import argparse
import json
import os
import platform
import resource
import sys
import time

from .config import SimulationConfig
from .constants import (
    AIR_DENSITY_20C_KG_M3,
    ION_DIFFUSION_NEGATIVE_CM2_S,
    ION_DIFFUSION_POSITIVE_CM2_S,
    ION_MOBILITY_NEGATIVE_CM2_VS,
    ION_MOBILITY_POSITIVE_CM2_VS,
)

# The scenario (run_markus_2mm.py:71-113): beam and chamber, Kanai carriers,
# dry air at 20 degC, reflecting wall, 2.6 separation times of clearance.
BEAM_KWARGS = dict(E_MeV_u=56.2, voltage_V=300.0, electrode_gap_cm=0.2, pulse_duration_s=540e-6,
                   repetition_rate_hz=50.0, n_pulses=1, rf_frequency_hz=26.26e6, seed=20260527)
WATER_TO_AIR = 0.891
DEFAULT_DOSE_RATE_WATER_GY_S = 10.0
CHAMBER_PHYSICS_KWARGS = dict(
    mu_positive_cm2_Vs=ION_MOBILITY_POSITIVE_CM2_VS, mu_negative_cm2_Vs=ION_MOBILITY_NEGATIVE_CM2_VS,
    D_positive_cm2_s=ION_DIFFUSION_POSITIVE_CM2_S, D_negative_cm2_s=ION_DIFFUSION_NEGATIVE_CM2_S,
    W_eV=33.0, air_density_kg_m3=AIR_DENSITY_20C_KG_M3, lateral_boundary="reflecting",
    n_clearance_separation_times=2.6, track_cutoff_sigmas=10.0,
)
# name: (sampled radius cm, buffer radius), 10 um voxels (run_markus_2mm.py:120-131)
GRID_TIERS = {
    "dev": (0.003, 3),
    "archive": (0.008, 3),
    "standard": (0.014, 3),
    "wide": (0.018, 3),
    "full_electrode": (0.265, 3),
}
DEFAULT_TIER = "archive"
# finite-column edge deficit of k_s, falling as 1/radius (run_markus_2mm.py:133-138)
EDGE_DEFICIT_UM = 1.512


def build_config(tier: str = DEFAULT_TIER, dose_rate_water_Gy_s: float = DEFAULT_DOSE_RATE_WATER_GY_S,
                 sampled_radius_cm=None, grid_size_um: float = 10.0) -> SimulationConfig:
    if tier not in GRID_TIERS:
        raise ValueError(f"Unknown tier {tier!r}; expected one of {sorted(GRID_TIERS)}.")
    radius, buffer_radius = GRID_TIERS[tier]
    return SimulationConfig(
        **BEAM_KWARGS, **CHAMBER_PHYSICS_KWARGS, dose_rate_Gy_s=dose_rate_water_Gy_s * WATER_TO_AIR,
        grid_size_um=grid_size_um, sampled_radius_cm=radius if sampled_radius_cm is None else sampled_radius_cm,
        buffer_radius=buffer_radius, no_z_electrode=5, memory_budget_fraction=None,
    )


def _gib(n: float) -> str:
    return f"{n / 2**30:.3f} GiB"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("tier", nargs="?", default=DEFAULT_TIER, choices=sorted(GRID_TIERS))
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"))
    ap.add_argument("--json", default=None, help="also write the result as JSON")
    ap.add_argument("--save-run", default=None, metavar="DIR",
                    help="write collected_charge.csv, track_density_xy.npy and run_meta.json to DIR")
    ap.add_argument("--sampled-radius-mm", type=float, default=None)
    ap.add_argument("--dose-rate-water-Gy-s", type=float, default=DEFAULT_DOSE_RATE_WATER_GY_S)
    ap.add_argument("--grid-um", type=float, default=10.0, help="voxel size (the tiers are quoted at 10 um)")
    ap.add_argument("--tiling", default=None, choices=("auto",), help="autotune the sweep tiling (cached)")
    sizing = ap.add_mutually_exclusive_group()
    sizing.add_argument("--dry-run", action="store_true", help="print the device memory sizing and exit")
    sizing.add_argument("--estimate-runtime-seconds", type=float, default=None, metavar="N",
                        help="run the real grid for ~N seconds, extrapolate, exit")
    args = ap.parse_args(argv)

    import ctypes

    from . import _lib
    from .solver_hip import params_from_config, run_simulation_hip, warmup_hip

    radius_cm = None if args.sampled_radius_mm is None else args.sampled_radius_mm / 10.0
    config = build_config(args.tier, args.dose_rate_water_Gy_s, radius_cm, args.grid_um)
    print(f"=== IFJ AIC-144, Markus 2 mm, macropulse, {args.dose_rate_water_Gy_s:g} Gy/s to water"
          f" -- '{args.tier}' grid ===")
    print(config.summary())
    device, free, total = _lib.device_info()
    print(f"Backend               : hip ({device}), {args.precision} fields")
    print()

    if args.dry_run:
        need = ctypes.c_size_t()
        params = params_from_config(config)
        _lib.check(_lib.open_library().pic_mem_required(ctypes.byref(params), _lib.PRECISIONS[args.precision],
                                                        ctypes.byref(need)), "pic_mem_required")
        tracks = config.number_of_tracks_per_pulse * config.n_pulses
        print("--- dry run: device memory ---")
        print(f"carrier arrays + scratch : {_gib(need.value)}")
        print(f"track centres            : {_gib(16 * tracks)} ({tracks:,} tracks)")
        print(f"free / total on device   : {_gib(free)} / {_gib(total)}")
        print("\nNo simulation was run (--dry-run).")
        return 0

    if args.estimate_runtime_seconds is not None:
        from .benchmark import estimate_full_runtime_empirical

        est = estimate_full_runtime_empirical(config, args.estimate_runtime_seconds, args.precision)
        print(f"--- empirical runtime estimate: hip backend, ~{args.estimate_runtime_seconds:g}s sample ---")
        print(f"steps measured            : {est['steps_measured']:,} / {est['total_time_steps']:,}")
        print(f"measured time for those   : {est['elapsed_measured_s']:.2f} s ({est['ms_per_step_measured']:.3f} ms/step)")
        if est["exact"]:
            print("\nThe whole run finished inside the budget -- the figure below is the real loop time.")
        print(f"estimated total loop time : {est['estimated_seconds']:.2f} s")
        print("\nNo full run was performed (--estimate-runtime-seconds).")
        return 0

    warmup_hip()
    t0 = time.perf_counter()
    result = run_simulation_hip(config, progress=True, precision=args.precision, tiling=args.tiling)
    elapsed = time.perf_counter() - t0
    peak_rss_kb = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
    radius_um = config.sampled_radius_cm * 1e4
    corrected = result.ks + EDGE_DEFICIT_UM / radius_um
    interior = (config.no_xy - 2) ** 2 * (config.no_z_with_buffer - 2)
    print(f"\nWall time (hip, {args.precision}): {elapsed:.2f} s "
          f"({interior * config.total_time_steps / elapsed:.3g} voxel-steps/s, track draw and upload included)")
    print(f"Collection efficiency f = {result.f_t[-1]:.4f}")
    print(f"Recombination correction k_s = 1/f = {result.ks:.4f}")
    print(f"Corrected for the finite-column edge deficit (+{EDGE_DEFICIT_UM / radius_um:.4f}): k_s -> {corrected:.4f}")

    if args.save_run:
        from .output import save_run_record

        paths = save_run_record(result, args.save_run)
        for key in ("csv", "track_density", "meta"):
            print(f"Wrote {paths[key]}")
    if args.json:
        with open(args.json, "w") as handle:
            json.dump({
                "tier": args.tier, "sampled_radius_cm": config.sampled_radius_cm, "backend": "hip",
                "precision": args.precision, "device": device,
                "dose_rate_water_Gy_s": args.dose_rate_water_Gy_s, "dose_rate_air_Gy_s": config.dose_rate_Gy_s,
                "hostname": platform.node(), "affinity_cpus": len(os.sched_getaffinity(0)),
                "wall_s": elapsed, "peak_rss_kb": peak_rss_kb,
                "f": float(result.f_t[-1]), "ks": float(result.ks), "ks_corrected": float(corrected),
                "no_xy": config.no_xy, "no_z_with_buffer": config.no_z_with_buffer,
                "total_time_steps": config.total_time_steps, "tracks_per_pulse": config.number_of_tracks_per_pulse,
            }, handle, indent=2)
        print(f"wrote {args.json}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
