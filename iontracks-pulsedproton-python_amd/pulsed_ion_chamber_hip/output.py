"""A run's per-time-step record on disk, in the reference's format.

The format is the reference's (pulsed_ion_chamber/output.py:63-198), kept so
that a record written here plots with its ``examples/ifj_aic144/plot.py``.
One directory holds three files:

* ``collected_charge.csv`` -- a header line and one line per time step with
  the columns of ``COLLECTED_CHARGE_COLUMNS``, each value ``%.12e``, comma
  separated, lines ending in CR LF (what Python's csv writer produces); every
  column except ``time`` (seconds, end of the step) is in ion pairs: the
  per-step sums of densities times the voxel volume;
* ``track_density_xy.npy`` -- the number of track centres per voxel column;
* ``run_meta.json`` -- the config scalars the plots read
  (``PLOT_CONFIG_FIELDS``) and the run's f, k_s, tracks per pulse and
  injected and recombined ion pairs.

The functions are written from that description: the table is built by
NumPy, the CSV written and read with ``numpy.savetxt`` / ``numpy.loadtxt``.
"""

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

# file name of each part of a record
_CSV, _DENSITY, _META = "collected_charge.csv", "track_density_xy.npy", "run_meta.json"
# Result attribute behind each column; every column but time is scaled to ion pairs
_SOURCE = {"time": "time_s", "n_positive": "n_positive", "n_negative": "n_negative",
           "injected_positive": "injected_positive", "injected_negative": "injected_negative",
           "recombination": "recombination"}


def _voxel_volume(result) -> float:
    return result.config.unit_length_cm**3


def collected_charge_table(result) -> dict:
    """The record's columns as arrays: ``time`` in seconds, the rest in ion pairs."""
    volume = _voxel_volume(result)
    table = {}
    for column in COLLECTED_CHARGE_COLUMNS:
        values = np.asarray(getattr(result, _SOURCE[column]), dtype=np.float64)
        table[column] = values if column == "time" else values * volume
    return table


def write_collected_charge_csv(result, path: Union[str, Path]) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    table = collected_charge_table(result)
    matrix = np.column_stack([table[c] for c in COLLECTED_CHARGE_COLUMNS])
    np.savetxt(path, matrix, fmt="%.12e", delimiter=",", newline="\r\n",
               header=",".join(COLLECTED_CHARGE_COLUMNS), comments="")
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


def _summary(result) -> dict:
    cfg = result.config
    volume = _voxel_volume(result)
    return {
        "f": float(result.f_t[-1]),
        "ks": float(result.ks),
        "tracks_per_pulse": int(cfg.number_of_tracks_per_pulse),
        "injected_ion_pairs": float(np.asarray(result.injected_positive).sum() * volume),
        "recombined_ion_pairs": float(np.asarray(result.recombination).sum() * volume),
    }


def save_run_record(result, directory: Union[str, Path]) -> dict:
    """Write the three files; returns their paths keyed csv, track_density, meta."""
    where = Path(directory)
    where.mkdir(parents=True, exist_ok=True)
    paths = {"csv": where / _CSV, "track_density": where / _DENSITY, "meta": where / _META}
    write_collected_charge_csv(result, paths["csv"])
    np.save(paths["track_density"], np.asarray(result.track_density_xy))
    meta = {name: getattr(result.config, name) for name in PLOT_CONFIG_FIELDS}
    meta.update(_summary(result))
    paths["meta"].write_text(json.dumps(meta, indent=2) + "\n")
    return paths


def load_run_record(directory: Union[str, Path]) -> RunRecord:
    """Read a record back. A missing part raises FileNotFoundError with its name."""
    where = Path(directory)
    missing = [name for name in (_CSV, _DENSITY, _META) if not (where / name).exists()]
    if missing:
        raise FileNotFoundError(
            f"{where} holds no complete run record: {', '.join(missing)} missing "
            "(save_run_record writes all three files)")
    meta = json.loads((where / _META).read_text())
    with (where / _CSV).open() as handle:
        header = handle.readline().strip().split(",")
    body = np.loadtxt(where / _CSV, delimiter=",", skiprows=1, ndmin=2)
    columns = dict(zip(header, body.T))
    return RunRecord(
        config=SimpleNamespace(**{name: meta[name] for name in PLOT_CONFIG_FIELDS}),
        table={name: np.array(columns[name]) for name in COLLECTED_CHARGE_COLUMNS},
        track_density_xy=np.load(where / _DENSITY, allow_pickle=False),
        meta=meta,
    )
