"""Autotuner for the sweep tiling: rows per tile, planes per block, values
per lane.

The library picks a tiling by a rule fitted to measurements on MI355X
(enough waves for the chip, then the tallest tile and longest chunk). A grid
far from the measured ones may prefer another shape, so this module measures:
it runs the first in-pulse steps of the configuration's own schedule under
each candidate tiling and under the library's own choice (HIP events around
every sweep launch), keeps the fastest, and caches it in a JSON file keyed by device, library build, grid,
precision and wall (a rebuilt library may offer other tilings). Every tiling gives the same field values -- each voxel is updated by
the same arithmetic -- so tuning changes speed only (and, within rounding,
the order of the scored partial sums).

The cache lives at $PIC_HIP_TUNING_CACHE, or
~/.cache/pulsed_ion_chamber_hip/tiling.json.
"""

import json
import os
from pathlib import Path
from typing import Iterable, Optional, Tuple

from . import _lib
from .pulses import draw_tracks
from .solver_hip import HipSimulation

Tiling = Tuple[int, int, int]  # rows per tile, planes per block, values per lane


def cache_path() -> Path:
    env = os.environ.get("PIC_HIP_TUNING_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "pulsed_ion_chamber_hip" / "tiling.json"


def cache_key(config, precision: str) -> str:
    device = _lib.device_info()[0]
    grid = f"{config.no_xy}x{config.no_xy}x{config.no_z_with_buffer}"
    return f"{device}|lib {_lib.library_digest()}|{grid}|{precision}|{config.lateral_boundary}"


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def candidates(precision: str) -> list:
    """Tile shapes the library instantiates: rows 1, 2, 4 (FP32: 8 with one
    value per lane, 1 to 3 with four; FP64: 1 and 2 with two values per lane),
    4 to 32 planes per block."""
    out = []
    lanes = (4, 2, 1) if precision == "fp32" else (2, 1)
    for v in lanes:
        if precision == "fp32":
            rows = {1: (1, 2, 4, 8), 2: (1, 2, 4), 4: (1, 2, 3)}[v]
        else:
            rows = {1: (1, 2, 4), 2: (1, 2)}[v]
        for tj in rows:
            for chunk in (4, 8, 16, 32):
                out.append((tj, chunk, v))
    return out


def cached_tiling(config, precision: str) -> Optional[Tiling]:
    entry = _load(cache_path()).get(cache_key(config, precision))
    return tuple(entry["tiling"]) if entry else None


def autotune(config, precision: str = "fp32", steps: int = 48, repeats: int = 2,
             tilings: Optional[Iterable[Tiling]] = None, use_cache: bool = True, refresh: bool = False) -> dict:
    """Measure the candidate tilings and return
    ``{"tiling": (rows, planes, values), "sweep_us": ..., "default_us": ..., "measured": {...}}``.
    With ``use_cache`` a cached result is returned without measuring, and a
    new result is stored; ``refresh`` measures even when an entry exists and
    replaces it."""
    key = cache_key(config, precision)
    path = cache_path()
    if use_cache and not refresh:
        entry = _load(path).get(key)
        if entry:
            return dict(entry, cached=True)
    schedule, xs, ys = draw_tracks(config)
    steps = max(1, min(int(steps), config.total_time_steps))
    measured = {}
    with HipSimulation(config, precision) as sim:
        sim.upload_tracks(schedule, xs, ys)

        def time_current():
            best = float("inf")
            for _ in range(repeats):
                sim.reset()
                sim.enable_timing(True)
                sim.run_steps(0, steps)
                t = sim.timing()
                # per step: a launch of the two-step pass covers two
                best = min(best, t["sweep_ms"] * 1e3 / steps)
            return best

        sim.set_tiling(0, 0, 0)
        sim.run_steps(0, min(steps, 8))  # load the kernels
        default_us = time_current()
        # the library's own choice competes too: it may take a shape outside
        # the candidate list (the short march's 1- and 2-plane chunks, the
        # one-round chunk lengths between the powers of two, 40 planes)
        t = sim.get_tiling()
        default_tiling = (t["rows_per_tile"], t["planes_per_chunk"], t["values_per_lane"])
        measured[default_tiling] = default_us
        for tiling in (tilings or candidates(precision)):
            try:
                sim.set_tiling(*tiling)
            except (ValueError, MemoryError):
                continue
            measured[tiling] = time_current()
    best = min(measured, key=measured.get)
    result = {"tiling": list(best), "sweep_us": measured[best], "default_us": default_us,
              "default_tiling": list(default_tiling),
              "steps": steps, "measured": {",".join(map(str, k)): v for k, v in measured.items()}}
    if use_cache:
        table = _load(path)
        table[key] = result
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(table, indent=1))
        tmp.replace(path)
    return dict(result, cached=False)
