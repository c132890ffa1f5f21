"""A run's per-time-step record on disk, in the reference's format.

The record is what the reference's plotting reads, so the format is kept:
pulsed_ion_chamber/output.py:63-198 writes, in one directory,

* ``collected_charge.csv`` -- columns time, n_positive, n_negative,
  injected_positive, injected_negative, recombination; every value in ion
  pairs (densities times the voxel volume), ``%.12e``;
* ``track_density_xy.npy`` -- track centres per voxel column;
* ``run_meta.json`` -- the config scalars the plots read plus f, k_s and
  totals.

:func:`save_run_record` / :func:`load_run_record` write and read it, so a
record made by the HIP backend plots with the reference's own
``examples/ifj_aic144/plot.py``.
"""

import csv
import json
from dataclasses import dataclass
from pathlib import Path
from types import SimpleNamespace
from typing import Union

import numpy as np

COLLECTED_CHARGE_COLUMNS = ("time", "n_positive", "n_negative", "injected_positive", "injected_negative",
                            "recombination")
# config scalars the reference's plots read (output.py:42-51)
PLOT_CONFIG_FIELDS = ("dt", "pulse_duration_s", "no_xy", "unit_length_cm", "inner_radius", "sampling_radius",
                      "chamber_fill_fraction", "mid_xy")


def collected_charge_table(result) -> dict:
    """Per-step record in ion pairs: densities summed over voxels times the
    voxel volume; ``time`` is the end of each step in seconds."""
    volume = result.config.unit_length_cm**3
    return {
        "time": np.asarray(result.time_s),
        "n_positive": result.n_positive * volume,
        "n_negative": result.n_negative * volume,
        "injected_positive": result.injected_positive * volume,
        "injected_negative": result.injected_negative * volume,
        "recombination": result.recombination * volume,
    }


def write_collected_charge_csv(result, path: Union[str, Path]) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    table = collected_charge_table(result)
    columns = [table[name] for name in COLLECTED_CHARGE_COLUMNS]
    with path.open("w", newline="") as handle:
        out = csv.writer(handle)
        out.writerow(COLLECTED_CHARGE_COLUMNS)
        for row in zip(*columns):
            out.writerow([f"{float(v):.12e}" for v in row])
    return path


def track_density_per_cm2(result) -> np.ndarray:
    """Track centres per cm^2 of each voxel column."""
    return result.track_density_xy / result.config.unit_length_cm**2


@dataclass
class RunRecord:
    config: SimpleNamespace  # only PLOT_CONFIG_FIELDS
    table: dict              # as collected_charge_table, in ion pairs
    track_density_xy: np.ndarray
    meta: dict


def save_run_record(result, directory: Union[str, Path]) -> dict:
    """Write the three files; returns their paths keyed csv, track_density, meta."""
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    cfg = result.config
    csv_path = write_collected_charge_csv(result, directory / "collected_charge.csv")
    density_path = directory / "track_density_xy.npy"
    np.save(density_path, np.asarray(result.track_density_xy))
    volume = cfg.unit_length_cm**3
    meta = {name: getattr(cfg, name) for name in PLOT_CONFIG_FIELDS}
    meta.update(
        f=float(result.f_t[-1]),
        ks=float(result.ks),
        tracks_per_pulse=int(cfg.number_of_tracks_per_pulse),
        injected_ion_pairs=float(np.sum(result.injected_positive) * volume),
        recombined_ion_pairs=float(np.sum(result.recombination) * volume),
    )
    meta_path = directory / "run_meta.json"
    meta_path.write_text(json.dumps(meta, indent=2) + "\n")
    return {"csv": csv_path, "track_density": density_path, "meta": meta_path}


def load_run_record(directory: Union[str, Path]) -> RunRecord:
    """Read what save_run_record wrote; FileNotFoundError names a missing file."""
    directory = Path(directory)
    paths = [directory / "collected_charge.csv", directory / "track_density_xy.npy", directory / "run_meta.json"]
    for path in paths:
        if not path.exists():
            raise FileNotFoundError(f"{path} not found -- is {directory} a directory written by save_run_record()?")
    meta = json.loads(paths[2].read_text())
    with paths[0].open(newline="") as handle:
        rows = list(csv.reader(handle))
    header, body = rows[0], rows[1:]
    cols = {name: np.array([float(r[q]) for r in body]) for q, name in enumerate(header)}
    return RunRecord(
        config=SimpleNamespace(**{name: meta[name] for name in PLOT_CONFIG_FIELDS}),
        table={name: cols[name] for name in COLLECTED_CHARGE_COLUMNS},
        track_density_xy=np.load(paths[1], allow_pickle=False),
        meta=meta,
    )
