"""Empirical runtime estimate for the HIP backend.

The reference's estimator (pulsed_ion_chamber/benchmark.py:127-216) runs the
real CPU backend on the real grid for a bounded wall time and extrapolates
linearly in the step count; it knows only the CPU backends. This is the same
idea on the GPU -- the real configuration, real allocation, real track upload,
real kernels -- with one refinement the GPU makes cheap: a sweep's cost does
not depend on the field values, so steps can be timed out of order. Half of
the budget times steps that deposit tracks (the start of the first pulse),
the other half steps that do not (the clearance after it), and each rate is
applied to its own share of the run. (The reference's linear extrapolation
from the first steps alone over-estimates by the deposition share, which on
the GPU is the only systematic error: there is no thermal clock drift of the
kind its docstring describes for laptops.)
"""

from time import perf_counter

import numpy as np

from .pulses import draw_tracks
from .solver_hip import HipSimulation, warmup_hip

# steps per timed launch batch: run_simulation_hip's own chunk, so that the
# sample pays the same per-chunk synchronisation as the run it estimates
# (on a launch-bound grid of ~7 us steps, 32-step chunks over-estimated by 17 %)
CHUNK = 256


def _time_steps(sim, first, last, budget_s):
    """Run steps from ``first`` (below ``last``) in chunks until the budget
    is spent or ``last`` is reached; returns (steps, seconds)."""
    sim.synchronize()
    t0 = perf_counter()
    done = 0
    while first + done < last:
        count = min(CHUNK, last - first - done)
        sim.run_steps(first + done, count)
        sim.synchronize()
        done += count
        if perf_counter() - t0 >= budget_s:
            break
    return done, perf_counter() - t0


def estimate_full_runtime_empirical(config, max_wall_s: float = 5.0, precision: str = "fp64", rng=None) -> dict:
    """Estimate the loop time of ``run_simulation_hip(config, precision=...)``
    from about ``max_wall_s`` seconds of the real run. Keys as the
    reference's estimator (steps_measured, total_time_steps,
    elapsed_measured_s, ms_per_step_measured, exact, estimated_seconds,
    estimated_hours), plus the two per-step costs and the setup time. If the
    whole run fits in the first half of the budget the result is that run's
    time (``exact``). Raises RuntimeError if the budget allowed no step at
    all (as the reference's does)."""
    warmup_hip()
    t0 = perf_counter()
    schedule, xs, ys = draw_tracks(config, rng)
    total = config.total_time_steps
    deposits = np.asarray(schedule[:total]) > 0
    n_dep = int(deposits.sum())
    n_plain = total - n_dep
    with HipSimulation(config, precision) as sim:
        sim.upload_tracks(schedule, xs, ys)
        setup_s = perf_counter() - t0
        # sample A: the run from its start (the first pulse: mostly depositing steps)
        a, sec_a = _time_steps(sim, 0, total, max_wall_s / 2)
        samples = [(int(deposits[:a].sum()), a - int(deposits[:a].sum()), sec_a)]
        exact = a == total
        if not exact:
            # sample B: from the first step without tracks at or after A's end
            # (the clearance after the pulse)
            rest = np.flatnonzero(~deposits[a:])
            b0 = a + int(rest[0]) if rest.size else a
            b, sec_b = _time_steps(sim, b0, total, max_wall_s / 2)
            samples.append((int(deposits[b0:b0 + b].sum()), b - int(deposits[b0:b0 + b].sum()), sec_b))
    steps = sum(d + p for d, p, _ in samples)
    elapsed = sum(t for _, _, t in samples)
    if steps == 0:
        raise RuntimeError(f"max_wall_s={max_wall_s:.3g} s was not enough for one step -- raise max_wall_s.")
    # per-step cost of a depositing and of a plain step: the two samples as
    # two equations d x + p y = t; where they do not separate the two kinds,
    # one rate for all steps
    x = y = elapsed / steps
    if len(samples) == 2:
        (d1, p1, t1), (d2, p2, t2) = samples
        det = d1 * p2 - d2 * p1
        if det != 0:
            xs_, ys_ = (t1 * p2 - t2 * p1) / det, (d1 * t2 - d2 * t1) / det
            if xs_ > 0 and ys_ > 0:
                x, y = xs_, ys_
    estimated = elapsed if exact else x * n_dep + y * n_plain
    return {
        "backend": "solver_hip",
        "precision": precision,
        "steps_measured": int(steps),
        "total_time_steps": int(total),
        "elapsed_measured_s": float(elapsed),
        "ms_per_step_measured": 1e3 * elapsed / steps,
        "ms_per_depositing_step": 1e3 * x,
        "ms_per_plain_step": 1e3 * y,
        "depositing_steps": n_dep,
        "exact": bool(exact),
        "estimated_seconds": float(estimated),
        "estimated_hours": float(estimated) / 3600.0,
        "setup_and_upload_s": float(setup_s),
    }
