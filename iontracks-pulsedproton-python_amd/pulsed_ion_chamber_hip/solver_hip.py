"""``run_simulation_hip``: the pulsed-ion-chamber time-stepping loop on an AMD
Instinct MI355X (gfx950), a drop-in sibling of the reference's
``run_simulation_numba`` / ``run_simulation_numba_parallel`` /
``run_simulation_cuda`` (pulsed_ion_chamber/solver_cuda.py:645-899).

Host side (here): draw the schedule and every track centre of the run with
the reference's RNG stream, upload them once, then let the library run the
steps in chunks. Device side (libpic_hip.so): per step a gather deposition,
the fused Lax-Wendroff sweep with its scored reduction, the lateral-ring
update and a compensated per-step reduction, all on one HIP stream with the
four carrier arrays resident in HBM. Only the (n_steps, 4) per-step record
and, optionally, the final fields come back.
"""

import ctypes
from time import perf_counter
from typing import Optional

import numpy as np

from . import _lib
from .constants import RECOMBINATION_ALPHA_CM3_S
from .pulses import draw_tracks
from .state import build_result, track_density_xy


def params_from_config(config, alpha_cm3_s: float = RECOMBINATION_ALPHA_CM3_S) -> "_lib.PicParams":
    """Flatten the attributes the loop reads (config.py contract) into the C
    struct. ``alpha_dt = alpha * dt`` as solver_numba.py:274 forms it."""
    (p_lat, p_zm, p_zp, p_cen), (n_lat, n_zm, n_zp, n_cen) = config.scheme_coefficients()
    h2 = config.unit_length_cm**2
    b2 = config.track_radius_cm**2
    return _lib.PicParams(
        no_xy=int(config.no_xy),
        nzb=int(config.no_z_with_buffer),
        nze=int(config.no_z_electrode),
        no_z=int(config.no_z),
        mid_xy=float(config.mid_xy),
        scoring_radius_sq=float(config.scoring_radius_sq),
        p_lat=p_lat, p_zm=p_zm, p_zp=p_zp, p_cen=p_cen,
        n_lat=n_lat, n_zm=n_zm, n_zp=n_zp, n_cen=n_cen,
        alpha_dt=alpha_cm3_s * config.dt,
        gaussian_factor=float(config.Gaussian_factor),
        h2_over_b2=h2 / b2,
        cutoff_voxels=float(config.track_cutoff_voxels),
        wall=_lib.PIC_WALL_REFLECTING if config.lateral_boundary == "reflecting" else _lib.PIC_WALL_ABSORBING,
        reserved=0,
    )


class HipSimulation:
    """One device state (four carrier arrays + scratch + stream) for a config.

    Kernel-level handle used by :func:`run_simulation_hip` and by tests that
    drive deposition and single sweeps directly, the way the reference's
    tests/test_charge_conservation.py drives its Numba kernels.
    """

    def __init__(self, config, precision: str = "fp64", params=None):
        if precision not in _lib.PRECISIONS:
            raise ValueError(f'precision must be "fp64" or "fp32", got {precision!r}.')
        self.L = _lib.open_library()
        _lib.device_info()
        self.config = config
        self.precision = precision
        # ``params`` (a PicParams) overrides the weights, coefficients and
        # deposition constants derived from ``config`` (the grid must agree)
        self.params = params if params is not None else params_from_config(config)
        handle = ctypes.c_void_p()
        _lib.check(self.L.pic_create(ctypes.byref(self.params), _lib.PRECISIONS[precision], ctypes.byref(handle)),
                   "pic_create")
        self.handle = handle
        self.shape = (config.no_xy, config.no_xy, config.no_z_with_buffer)

    def close(self):
        if getattr(self, "handle", None):
            self.L.pic_destroy(self.handle)
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

    # --- run level -------------------------------------------------------
    def upload_tracks(self, schedule: np.ndarray, xs: np.ndarray, ys: np.ndarray) -> None:
        offsets = np.zeros(len(schedule) + 1, dtype=np.int64)
        np.cumsum(schedule, out=offsets[1:])
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        if len(xs) != offsets[-1] or len(ys) != offsets[-1]:
            raise ValueError("xs/ys length does not match the schedule total")
        _lib.check(self.L.pic_upload_tracks(self.handle, _lib.dp(xs), _lib.dp(ys), len(xs), _lib.ip(offsets),
                                            len(schedule)), "pic_upload_tracks")
        self.n_steps = len(schedule)

    def run_pulse(self, schedule: np.ndarray, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
        """Upload, run every step and return the (n_steps, 4) record in one call (pic_run_pulse)."""
        offsets = np.zeros(len(schedule) + 1, dtype=np.int64)
        np.cumsum(schedule, out=offsets[1:])
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        if len(xs) != offsets[-1] or len(ys) != offsets[-1]:
            raise ValueError("xs/ys length does not match the schedule total")
        out = np.zeros((len(schedule), 4))
        _lib.check(self.L.pic_run_pulse(self.handle, _lib.dp(xs), _lib.dp(ys), _lib.ip(offsets), len(schedule),
                                        _lib.dp(out)), "pic_run_pulse")
        self.n_steps = len(schedule)
        return out

    def run_pulse_static(self, schedule: np.ndarray, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
        """As :meth:`run_pulse`, through the entry point that allocates nothing
        (pic_run_pulse_static: inputs staged in the library's static 64 MiB
        device array; the handle's uploaded tracks are left untouched)."""
        offsets = np.zeros(len(schedule) + 1, dtype=np.int64)
        np.cumsum(schedule, out=offsets[1:])
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        if len(xs) != offsets[-1] or len(ys) != offsets[-1]:
            raise ValueError("xs/ys length does not match the schedule total")
        out = np.zeros((len(schedule), 4))
        _lib.check(self.L.pic_run_pulse_static(self.handle, _lib.dp(xs), _lib.dp(ys), _lib.ip(offsets),
                                               len(schedule), _lib.dp(out)), "pic_run_pulse_static")
        return out

    def run_steps(self, first: int, count: int) -> None:
        _lib.check(self.L.pic_run_steps(self.handle, int(first), int(count)), "pic_run_steps")

    def read_per_step(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = self.n_steps - first if count is None else count
        out = np.zeros((count, 4))
        _lib.check(self.L.pic_read_per_step(self.handle, int(first), int(count), _lib.dp(out)), "pic_read_per_step")
        return out

    def reset(self) -> None:
        _lib.check(self.L.pic_reset(self.handle), "pic_reset")

    def synchronize(self) -> None:
        _lib.check(self.L.pic_synchronize(self.handle), "pic_synchronize")

    def set_gather_mode(self, mode: str = "auto") -> None:
        """How deposition finds the tracks reaching each tile: "auto", "scan"
        (every tile tests every track) or "binned" (tracks binned by tile on
        the device first). Same results bit for bit."""
        if mode not in _lib.GATHER_MODES:
            raise ValueError(f'gather mode must be "auto", "scan" or "binned", got {mode!r}.')
        _lib.check(self.L.pic_set_gather_mode(self.handle, _lib.GATHER_MODES[mode]), "pic_set_gather_mode")

    def gather_binned(self) -> bool:
        """Whether the uploaded tracks are binned by tile."""
        v = ctypes.c_int()
        _lib.check(self.L.pic_get_gather_mode(self.handle, ctypes.byref(v)), "pic_get_gather_mode")
        return bool(v.value)

    def set_tiling(self, rows_per_tile: int = 0, planes_per_chunk: int = 0, values_per_lane: int = 0) -> None:
        _lib.check(self.L.pic_set_tiling(self.handle, int(rows_per_tile), int(planes_per_chunk), int(values_per_lane)),
                   "pic_set_tiling")

    def set_steps_per_pass(self, steps: int = 0, planes_per_chunk: int = 0) -> None:
        """1 or 2 time steps per sweep launch (2: FP32 fields, either
        lateral wall; fields bit-identical to single steps); 0 for the library's
        choice. ``planes_per_chunk`` sets the two-step pass's planes per block."""
        _lib.check(self.L.pic_set_steps_per_pass(self.handle, int(steps), int(planes_per_chunk)),
                   "pic_set_steps_per_pass")

    def get_tiling(self) -> dict:
        """The tiling the next sweep launch uses (the library's choice for every
        knob set_tiling left at 0)."""
        v = [ctypes.c_int() for _ in range(3)]
        _lib.check(self.L.pic_get_tiling(self.handle, *[ctypes.byref(x) for x in v]), "pic_get_tiling")
        return dict(rows_per_tile=v[0].value, planes_per_chunk=v[1].value, values_per_lane=v[2].value)

    def enable_timing(self, on: bool = True, group: int = 1) -> None:
        """Time kernel launches with HIP events: each launch, or with ``group``
        > 1 runs of up to ``group`` back-to-back sweep launches per event pair
        (every sweep is still timed; the pairs cost the stream less)."""
        _lib.check(self.L.pic_enable_timing(self.handle, max(1, int(group)) if on else 0), "pic_enable_timing")

    def set_graphs(self, on: bool = True) -> None:
        """Replay run_steps' launches through graphs, or launch them one by one (the default)."""
        _lib.check(self.L.pic_set_graphs(self.handle, 1 if on else 0), "pic_set_graphs")

    def timing(self) -> dict:
        sweep_ms = ctypes.c_double()
        dep_ms = ctypes.c_double()
        other = ctypes.c_double()
        ns = ctypes.c_int64()
        nd = ctypes.c_int64()
        _lib.check(self.L.pic_get_timing(self.handle, ctypes.byref(sweep_ms), ctypes.byref(ns), ctypes.byref(dep_ms),
                                         ctypes.byref(nd), ctypes.byref(other)), "pic_get_timing")
        return dict(sweep_ms=sweep_ms.value, sweep_launches=ns.value, deposit_ms=dep_ms.value,
                    deposit_launches=nd.value, ring_reduce_ms=other.value)

    # --- kernel level ----------------------------------------------------
    def deposit(self, xs: np.ndarray, ys: np.ndarray) -> float:
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        out = ctypes.c_double()
        _lib.check(self.L.pic_deposit(self.handle, _lib.dp(xs), _lib.dp(ys), len(xs), ctypes.byref(out)),
                   "pic_deposit")
        return out.value

    def step(self):
        out = np.zeros(3)
        _lib.check(self.L.pic_step(self.handle, _lib.dp(out)), "pic_step")
        return tuple(float(v) for v in out)

    def read_fields(self):
        pos = np.empty(self.shape)
        neg = np.empty(self.shape)
        _lib.check(self.L.pic_read_fields(self.handle, _lib.dp(pos), _lib.dp(neg)), "pic_read_fields")
        return pos, neg

    def write_fields(self, pos: np.ndarray, neg: np.ndarray) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.shape)
        neg = np.ascontiguousarray(neg, dtype=np.float64).reshape(self.shape)
        _lib.check(self.L.pic_write_fields(self.handle, _lib.dp(pos), _lib.dp(neg)), "pic_write_fields")


def warmup_hip() -> None:
    """Load the library and run every kernel once on a 4x4x4 grid, so a later
    timed call measures execution only (the role warmup_cuda plays,
    solver_cuda.py:394-417)."""
    from types import SimpleNamespace

    tiny = SimpleNamespace(
        no_xy=4, no_z_with_buffer=4, no_z_electrode=1, no_z=2, mid_xy=2.0, scoring_radius_sq=1.0,
        unit_length_cm=1.0, track_radius_cm=1.0, Gaussian_factor=1.0, track_cutoff_voxels=4.0, dt=1e-9,
        lateral_boundary="absorbing",
        scheme_coefficients=lambda: ((0.01, 0.01, 0.01, 0.94), (0.01, 0.01, 0.01, 0.94)),
    )
    for precision in ("fp64", "fp32"):
        with HipSimulation(tiny, precision) as sim:
            sim.deposit(np.array([2.0]), np.array([2.0]))
            sim.step()
            sim.upload_tracks(np.array([1, 0], dtype=np.int64), np.array([2.0]), np.array([2.0]))
            sim.run_steps(0, 2)
            sim.synchronize()


def run_simulation_hip(
    config,
    rng: Optional[np.random.Generator] = None,
    progress: bool = True,
    max_wall_s: Optional[float] = None,
    max_steps: Optional[int] = None,
    return_fields: bool = True,
    precision: str = "fp64",
    chunk_steps: int = 256,
    tiling=None,
):
    """Simulate a pulse train on the GPU and return the reference's Result.

    Same physics and RNG stream as the reference backends: the schedule and
    track centres are drawn on the host by the same generator calls
    (``default_rng(config.seed)`` unless ``rng`` is given), so
    ``track_density_xy`` is identical and the fields agree with
    ``run_simulation_numba`` to rounding.

    ``precision="fp64"`` stores and sweeps FP64 fields; ``"fp32"`` stores and
    sweeps FP32 fields (half the HBM traffic) while deposition, the
    recombination product and every scored sum stay FP64.

    ``tiling`` picks the sweep's tile shape: None for the library's rule,
    ``"auto"`` for the autotuned shape (measured once per device, grid,
    precision and wall, then cached; see :mod:`.tuning`), or
    ``(rows_per_tile, planes_per_chunk, values_per_lane)``. Every tiling
    gives the same fields.

    ``max_wall_s`` / ``max_steps`` stop early exactly as in
    run_simulation_cuda (solver_cuda.py:664-671): ``ks`` is NaN and
    ``steps_completed`` / ``loop_elapsed_s`` are set. ``return_fields=False``
    returns empty ``(0, 0, 0)`` fields.
    """
    if precision not in _lib.PRECISIONS:
        raise ValueError(f'precision must be "fp64" or "fp32", got {precision!r}.')
    if max_steps is not None and max_steps < 0:
        raise ValueError(f"max_steps must be non-negative, got {max_steps!r}.")
    _lib.open_library()
    _lib.device_info()

    auto = tiling == "auto"
    if auto:
        from .tuning import autotune

        tiling = tuple(autotune(config, precision)["tiling"])
    elif tiling is not None and (isinstance(tiling, str) or len(tiling) != 3):
        raise ValueError(f'tiling must be None, "auto" or (rows, planes, values), got {tiling!r}.')
    schedule, xs, ys = draw_tracks(config, rng)
    total = config.total_time_steps
    last = total if max_steps is None else min(total, int(max_steps))
    with HipSimulation(config, precision) as sim:
        if tiling is not None:
            try:
                sim.set_tiling(*tiling)
            except ValueError:
                if not auto:
                    raise
                # a cached tiling this library no longer accepts: measure again
                from .tuning import autotune

                sim.set_tiling(*autotune(config, precision, refresh=True)["tiling"])
        sim.upload_tracks(schedule, xs, ys)
        report_every = max(1, total // 20)
        loop_t0 = perf_counter()
        done = 0
        next_report = 0
        while done < last:
            count = min(chunk_steps, last - done)
            sim.run_steps(done, count)
            done += count
            if max_wall_s is not None or (progress and done > next_report):
                sim.synchronize()
            if progress and done > next_report:
                print(f"  step {done}/{total}")
                next_report += report_every * max(1, -(-chunk_steps // report_every))
            if max_wall_s is not None and perf_counter() - loop_t0 >= max_wall_s:
                break
        sim.synchronize()
        loop_elapsed_s = perf_counter() - loop_t0
        per_step = np.zeros((total, 4))
        per_step[:done] = sim.read_per_step(0, done)
        if return_fields:
            positive, negative = sim.read_fields()
        else:
            positive = np.zeros((0, 0, 0))
            negative = np.zeros((0, 0, 0))
    used = int(schedule[:done].sum())
    result = build_result(config, per_step, done, positive, negative, track_density_xy(config, xs[:used], ys[:used]),
                          truncated=done < total)
    if max_wall_s is not None or max_steps is not None:
        result.steps_completed = done
        result.loop_elapsed_s = loop_elapsed_s
    return result
