"""Many configurations in one device stream: a voltage / dose-rate / seed
sweep (a saturation curve) run as one batch.

The reference runs a parameter study as independent processes, one
configuration each, every one executing the loop of
solver_numba_parallel.py:526-723 (docs/PERFORMANCE.md:211-218). Here each
configuration keeps its own device state (a :class:`HipSimulation` with its
own arrays, tracks and per-step record), and the library advances all of them
together: per global time step one gather launch per configuration and batch
of steps, and ONE sweep launch whose blocks cover every configuration still
running. Configurations with different grids and step counts mix freely;
each stops at its own last step. The numbers are those of running every
configuration on its own with :func:`run_simulation_hip` with the same
tiling: the same kernel body, the same arithmetic, the same order of every
sum (a different tiling changes only the order of the scored partial sums).
"""

import ctypes
from time import perf_counter
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .pulses import draw_tracks
from .solver_hip import HipSimulation
from .state import build_result, track_density_xy


class HipBatch:
    """A batch over existing handles (which it does not own)."""

    def __init__(self, sims: Sequence[HipSimulation]):
        if not sims:
            raise ValueError("a batch needs at least one simulation")
        self.L = _lib.open_library()
        self.sims = list(sims)
        arr = (ctypes.c_void_p * len(sims))(*[s.handle.value for s in sims])
        handle = ctypes.c_void_p()
        _lib.check(self.L.pic_batch_create(arr, len(sims), ctypes.byref(handle)), "pic_batch_create")
        self.handle = handle

    def close(self):
        if getattr(self, "handle", None):
            self.L.pic_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def longest(self) -> int:
        return max(s.n_steps for s in self.sims)

    def run(self, first: int = 0, count: Optional[int] = None) -> None:
        count = self.longest - first if count is None else count
        _lib.check(self.L.pic_batch_run(self.handle, int(first), int(count)), "pic_batch_run")

    def synchronize(self) -> None:
        _lib.check(self.L.pic_batch_synchronize(self.handle), "pic_batch_synchronize")

    def enable_timing(self, on: bool = True) -> None:
        _lib.check(self.L.pic_batch_enable_timing(self.handle, 1 if on else 0), "pic_batch_enable_timing")

    def set_graphs(self, on: bool = True) -> None:
        """Replay each segment's launches through graphs, or launch them one by one (the default)."""
        _lib.check(self.L.pic_batch_set_graphs(self.handle, 1 if on else 0), "pic_batch_set_graphs")

    def timing(self) -> dict:
        sweep_ms = ctypes.c_double()
        gather_ms = ctypes.c_double()
        ns = ctypes.c_int64()
        ng = ctypes.c_int64()
        _lib.check(self.L.pic_batch_get_timing(self.handle, ctypes.byref(sweep_ms), ctypes.byref(ns),
                                               ctypes.byref(gather_ms), ctypes.byref(ng)), "pic_batch_get_timing")
        return dict(sweep_ms=sweep_ms.value, sweep_launches=ns.value, gather_ms=gather_ms.value,
                    gather_launches=ng.value)

    def set_tiling(self, rows_per_tile: int = 0, planes_per_chunk: int = 0) -> None:
        _lib.check(self.L.pic_batch_set_tiling(self.handle, int(rows_per_tile), int(planes_per_chunk)),
                   "pic_batch_set_tiling")

    def tiling(self) -> dict:
        v = [ctypes.c_int() for _ in range(4)]
        _lib.check(self.L.pic_batch_get_tiling(self.handle, *[ctypes.byref(x) for x in v]), "pic_batch_get_tiling")
        return dict(rows_per_tile=v[0].value, planes_per_chunk=v[1].value, values_per_lane=v[2].value,
                    windows_per_block=v[3].value)


def run_batch_hip(
    configs: Sequence,
    precision: str = "fp32",
    rngs: Optional[Sequence[np.random.Generator]] = None,
    return_fields: bool = False,
    tiling: Optional[tuple] = None,
    timings: Optional[dict] = None,
):
    """Run every configuration to its last step in one batch and return one
    reference ``Result`` per configuration, in the given order.

    Each configuration draws its schedule and track centres exactly as
    :func:`run_simulation_hip` does (``default_rng(config.seed)`` unless
    ``rngs`` supplies a generator per configuration), so the results are
    those of running each configuration on its own. All configurations must
    share ``precision`` and the lateral boundary. ``tiling`` =
    ``(rows_per_tile, planes_per_chunk)`` overrides the automatic sweep
    tiling. If ``timings`` is a dict it receives the wall seconds of track
    drawing, device set-up and the batched loop.
    """
    if precision not in _lib.PRECISIONS:
        raise ValueError(f'precision must be "fp64" or "fp32", got {precision!r}.')
    if rngs is not None and len(rngs) != len(configs):
        raise ValueError("rngs must give one generator per configuration")
    walls = {c.lateral_boundary for c in configs}
    if len(walls) > 1:
        raise ValueError(f"configurations of one batch share a lateral boundary, got {sorted(walls)}")
    _lib.open_library()
    _lib.device_info()
    t0 = perf_counter()
    draws = [draw_tracks(c, None if rngs is None else rngs[q]) for q, c in enumerate(configs)]
    t1 = perf_counter()
    sims = []
    try:
        for c, (schedule, xs, ys) in zip(configs, draws):
            sim = HipSimulation(c, precision)
            sims.append(sim)
            sim.upload_tracks(schedule, xs, ys)
        with HipBatch(sims) as batch:
            if tiling is not None:
                batch.set_tiling(*tiling)
            batch.synchronize()
            t2 = perf_counter()
            batch.run()
            batch.synchronize()
            t3 = perf_counter()
        results = []
        for c, sim, (schedule, xs, ys) in zip(configs, sims, draws):
            per_step = sim.read_per_step(0, c.total_time_steps)
            if return_fields:
                positive, negative = sim.read_fields()
            else:
                positive = negative = np.zeros((0, 0, 0))
            results.append(build_result(c, per_step, c.total_time_steps, positive, negative,
                                        track_density_xy(c, xs, ys), truncated=False))
    finally:
        for sim in sims:
            sim.close()
    if timings is not None:
        timings.update(draw_s=t1 - t0, setup_s=t2 - t1, loop_s=t3 - t2)
    return results
