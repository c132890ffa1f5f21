"""Write tests/golden/published.json: the reference's own published numbers for
the time-stepping path, as data (no code of the reference is copied or run).

Sources, relative to the reference checkout:
  profiling/data/gpu_athena/ladder_10um.json          10 um ladder: grid, tracks, steps, k_s
  profiling/data/gpu_athena/gpu_5um_full_electrode.json and cpu_5um_full_electrode.json
  profiling/data/gh200/convergence.json               r 0.05 cm at 10/5/2 um: k_s
  profiling/data/gh200/resolution.json                200-step prefix f at 10/5/2/1 um
  profiling/data/gh200/full_electrode_2.0um.json      500-step prefix f of the 2 um full electrode
  examples/ifj_aic144/README.md (tier table)          dev/archive/standard/wide/full grids and tracks
  examples/ifj_aic144/results/full_electrode/collected_charge.csv  time column (dt)
  web/core/src/theory.rs:76-82                        Jaffe k_s golden for the single-track test scenario
  profiling/data/jaffe_validation.json                Jaffe k_s as evaluated by the reference's Python

Run from the repository root with the reference checkout at /root/reference.
"""

import csv
import json
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "published.json"


def rows(path):
    return json.loads((REF / path).read_text())["rows"]


def main():
    # wall times: the whole run_simulation_* call, FP64 fields (the GPU rows
    # on one A100-SXM4-40GB or one GH200, the CPU rows on 32 cores of an
    # Athena node, docs/GPU.md:22-30 and docs/BENCHMARKS-HELIOS-GH200.md)
    a100, gh200, cpu32 = "A100-SXM4-40GB (Numba-CUDA)", "GH200 (Numba-CUDA, device memory)", "32 CPU cores (batched)"
    out = {"aic_runs": [], "prefix_f": [], "tiers": [], "sources": __doc__}
    for row in rows("profiling/data/gpu_athena/ladder_10um.json"):
        out["aic_runs"].append({k: row[k] for k in ("sampled_radius_cm", "grid_size_um", "no_xy", "no_z_with_buffer",
                                                    "tracks_per_pulse", "total_time_steps", "cpu_ks", "gpu_ks")}
                               | {"gpu_wall_s": row["gpu_wall_s"], "gpu": a100, "cpu_wall_s": row["cpu_wall_s"],
                                  "cpu": cpu32})
    gpu = rows("profiling/data/gpu_athena/gpu_5um_full_electrode.json")[0]
    cpu = rows("profiling/data/gpu_athena/cpu_5um_full_electrode.json")[0]
    out["aic_runs"].append({k: gpu[k] for k in ("sampled_radius_cm", "grid_size_um", "no_xy", "no_z_with_buffer",
                                                "tracks_per_pulse", "total_time_steps", "gpu_ks")}
                           | {"cpu_ks": cpu["cpu_ks"], "gpu_wall_s": gpu["gpu_wall_s"], "gpu": a100,
                              "cpu_wall_s": cpu["cpu_wall_s"], "cpu": cpu32})
    for row in rows("profiling/data/gh200/convergence.json"):
        out["aic_runs"].append({"sampled_radius_cm": row["sampled_radius_cm"], "grid_size_um": row["grid_size_um"],
                                "no_xy": row["no_xy"], "no_z_with_buffer": row["no_z_with_buffer"],
                                "total_time_steps": row["total_time_steps"], "gpu_ks": row["ks"],
                                "gpu_wall_s": row["wall_s"], "gpu": gh200})
    for row in rows("profiling/data/gh200/resolution.json"):
        out["prefix_f"].append({"sampled_radius_cm": row["sampled_radius_cm"], "grid_size_um": row["grid_size_um"],
                                "steps": row["steps"], "f_end": row["f_end"], "no_xy": row["no_xy"],
                                "no_z_with_buffer": row["no_z_with_buffer"], "gpu_wall_s": row["wall_s"],
                                "gpu": gh200})
    for row in rows("profiling/data/gh200/full_electrode_2.0um.json"):
        out["prefix_f"].append({"sampled_radius_cm": row["sampled_radius_cm"], "grid_size_um": row["grid_size_um"],
                                "steps": row["steps"], "f_end": row["f_end"], "no_xy": row["no_xy"],
                                "no_z_with_buffer": row["no_z_with_buffer"],
                                "total_time_steps": row["total_time_steps"], "gpu_wall_s": row["wall_s"],
                                "gpu": "GH200 (Numba-CUDA, managed memory: the fields exceed its HBM)"})
    # examples/ifj_aic144/README.md tier table (name, radius, grid, tracks, rounded k_s)
    out["tiers"] = [
        {"tier": "dev", "sampled_radius_cm": 0.003, "no_xy": 12, "nzb": 210, "tracks": 3157, "ks_4dp": 1.0580},
        {"tier": "archive", "sampled_radius_cm": 0.008, "no_xy": 22, "nzb": 210, "tracks": 22447, "ks_4dp": 1.0929},
        {"tier": "standard", "sampled_radius_cm": 0.014, "no_xy": 34, "nzb": 210, "tracks": 68744, "ks_4dp": 1.1011},
        {"tier": "wide", "sampled_radius_cm": 0.018, "no_xy": 42, "nzb": 210, "tracks": 113638, "ks_4dp": 1.1035},
        {"tier": "full_electrode", "sampled_radius_cm": 0.265, "no_xy": 536, "nzb": 210, "tracks": 24630400,
         "ks_4dp": 1.1111},
    ]
    with open(REF / "examples/ifj_aic144/results/full_electrode/collected_charge.csv") as handle:
        reader = csv.reader(handle)
        next(reader)
        first = next(reader)
    out["full_electrode_10um_first_time_s"] = float(first[0])
    out["full_electrode_10um_steps"] = 2194
    out["jaffe_ks_single_track_golden"] = 1.0012341579794355
    out["jaffe_ks_validation_json"] = json.loads((REF / "profiling/data/jaffe_validation.json").read_text())[0]["ks_jaffe"]
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
