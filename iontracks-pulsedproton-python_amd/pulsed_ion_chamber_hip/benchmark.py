"""Empirical runtime estimate for the HIP backend.

The reference's estimator (pulsed_ion_chamber/benchmark.py:127-216) runs the
real CPU backend on the real grid for a bounded wall time and extrapolates;
it knows only the CPU backends. This is the same idea on the GPU: run the
real configuration (real allocation, real track upload, real kernels) for
about ``max_wall_s`` seconds and scale by the step count. The in-pulse steps
come first and cost a little more than the clearance steps (they deposit),
so the extrapolation slightly overestimates a run whose pulse is a small
part of it.
"""

from time import perf_counter

from .solver_hip import run_simulation_hip, warmup_hip


def estimate_full_runtime_empirical(config, max_wall_s: float = 5.0, precision: str = "fp64") -> dict:
    """Same keys as the reference's estimator: steps_measured,
    total_time_steps, elapsed_measured_s, ms_per_step_measured, exact,
    estimated_seconds, estimated_hours."""
    warmup_hip()
    t0 = perf_counter()
    result = run_simulation_hip(config, progress=False, max_wall_s=max_wall_s, return_fields=False,
                                precision=precision)
    wall = perf_counter() - t0
    total = config.total_time_steps
    steps = result.steps_completed
    elapsed = result.loop_elapsed_s
    exact = steps >= total
    estimated = elapsed if exact else elapsed * total / max(1, steps)
    return {
        "steps_measured": int(steps),
        "total_time_steps": int(total),
        "elapsed_measured_s": float(elapsed),
        "ms_per_step_measured": 1e3 * elapsed / max(1, steps),
        "exact": bool(exact),
        "estimated_seconds": float(estimated),
        "estimated_hours": float(estimated) / 3600.0,
        "setup_and_upload_s": float(wall - elapsed),
    }
