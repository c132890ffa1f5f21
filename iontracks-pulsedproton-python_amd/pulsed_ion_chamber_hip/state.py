"""The run record returned by every backend, and the per-step bookkeeping that
fills it (pulsed_ion_chamber/state.py:24-110)."""

from dataclasses import dataclass

import numpy as np


@dataclass
class Result:
    """Field names and meaning as state.py:24-48. Densities are cm^-3; the
    per-step series are density sums over the scored voxels."""

    config: object
    time_s: np.ndarray
    f_t: np.ndarray
    ks: float
    positive_array: np.ndarray
    negative_array: np.ndarray
    n_positive: np.ndarray
    n_negative: np.ndarray
    injected_positive: np.ndarray
    injected_negative: np.ndarray
    recombination: np.ndarray
    track_density_xy: np.ndarray


def track_density_xy(config, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
    """Histogram of track centres per voxel column, floats, via ``int()`` floors
    (state.py:71-84)."""
    flat = xs.astype(np.int64) * config.no_xy + ys.astype(np.int64)
    counts = np.bincount(flat, minlength=config.no_xy * config.no_xy)
    return counts.reshape(config.no_xy, config.no_xy).astype(np.float64)


def collection_efficiency(injected: np.ndarray, recombined: np.ndarray, n_steps: int) -> np.ndarray:
    """Running ``f_t`` exactly as the reference loops compute it
    (solver_numba.py:279-350): plain sequential FP64 sums of the per-step
    injected and recombined charge, 1.0 while nothing has been injected.
    Entries past ``n_steps`` keep their initial 1.0."""
    f_t = np.ones(len(injected))
    total_in = 0.0
    total_rec = 0.0
    inj = injected.tolist()
    rec = recombined.tolist()
    for step in range(n_steps):
        total_in += inj[step]
        total_rec += rec[step]
        f_t[step] = 1.0 if total_in == 0.0 else (total_in - total_rec) / total_in
    return f_t


def build_result(config, per_step: np.ndarray, steps_completed: int, positive, negative, tracks_xy,
                 truncated: bool) -> Result:
    """Assemble a :class:`Result` from a ``(total_time_steps, 4)`` array of
    per-step ``(injected, recombined, total_positive, total_negative)``."""
    injected = np.ascontiguousarray(per_step[:, 0])
    recombined = np.ascontiguousarray(per_step[:, 1])
    f_t = collection_efficiency(injected, recombined, steps_completed)
    time_s = (np.arange(config.total_time_steps) + 1) * config.dt
    ks = float("nan") if truncated else 1.0 / f_t[-1]
    return Result(
        config=config,
        time_s=time_s,
        f_t=f_t,
        ks=ks,
        positive_array=positive,
        negative_array=negative,
        n_positive=np.ascontiguousarray(per_step[:, 2]),
        n_negative=np.ascontiguousarray(per_step[:, 3]),
        injected_positive=injected,
        injected_negative=injected.copy(),
        recombination=recombined,
        track_density_xy=tracks_xy,
    )
