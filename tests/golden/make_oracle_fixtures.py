"""Write tests/golden/oracle_runs.json: the CPU oracle's results for the
benchmark configurations whose whole runs take the oracle too long to repeat
inside a GPU test (the mid-size 5 um solve, 5.8e10 voxel-steps) or that are
cheap to keep anyway (the README example, the 32-point saturation curve).

The oracle (oracle/pic_oracle.c, FP64) is itself pinned bit for bit to the
reference's published k_s values (tests/test_oracle_pins.py), so these are
the reference's numbers for configurations the reference never published.

Stored per run: k_s, f_t every 16 steps (and the last step), the summed
injected and recombined charge, and checksums of the final fields (sum,
sum of squares and maximum of each species). The numpy version is recorded
because the track draw depends on numpy's PCG64 stream.

Run from the repository root:  python tests/golden/make_oracle_fixtures.py
"""

import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
for p in (str(ROOT / "iontracks-pulsedproton-python_amd"), str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from configs import readme_config, saturation_curve_configs  # noqa: E402
from oracle import batched  # noqa: E402

OUT = Path(__file__).resolve().parent / "oracle_runs.json"
F_EVERY = 16


def field_checksums(a):
    return {"sum": float(a.sum()), "sum_sq": float(np.square(a).sum()), "max": float(a.max())}


def summary(cfg, res, fields=True):
    n = cfg.total_time_steps
    idx = list(range(F_EVERY - 1, n, F_EVERY))
    if idx[-1] != n - 1:
        idx.append(n - 1)
    out = {
        "grid": [cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer],
        "total_time_steps": n,
        "tracks": int(res.track_density_xy.sum()),
        "ks": float(res.ks),
        "f_t_index": idx,
        "f_t": [float(res.f_t[i]) for i in idx],
        "sum_injected": float(np.sum(res.injected_positive)),
        "sum_recombined": float(np.sum(res.recombination)),
    }
    if fields:
        out["positive"] = field_checksums(res.positive_array)
        out["negative"] = field_checksums(res.negative_array)
    return out


def main():
    out = {"numpy": np.__version__, "generator": "oracle/pic_oracle.c via oracle.batched.run_batched, FP64"}
    t0 = time.perf_counter()
    c0 = readme_config(10.0, 0.008)
    out["readme_10um_r0.008cm"] = summary(c0, batched.run_batched(c0))
    print(f"README example done ({time.perf_counter() - t0:.1f} s)", flush=True)
    t0 = time.perf_counter()
    c1 = readme_config(5.0, 0.05)
    out["mid_5um_r0.05cm"] = summary(c1, batched.run_batched(c1))
    print(f"mid-size solve done ({time.perf_counter() - t0:.1f} s)", flush=True)
    t0 = time.perf_counter()
    curve = []
    for cfg in saturation_curve_configs():
        res = batched.run_batched(cfg, fields=False)
        curve.append({"voltage_V": cfg.voltage_V, "dose_rate_Gy_s": cfg.dose_rate_Gy_s,
                      "grid_size_um": cfg.grid_size_um, "sampled_radius_cm": cfg.sampled_radius_cm,
                      **summary(cfg, res, fields=False)})
    out["saturation_curve"] = curve
    print(f"saturation curve done ({time.perf_counter() - t0:.1f} s)", flush=True)
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
