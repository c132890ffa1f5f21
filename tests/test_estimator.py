"""The GPU runtime estimator against the runs it estimates.

The reference's estimator (pulsed_ion_chamber/benchmark.py:127-216) is used
to size runs before committing to them; here the estimate from about a fifth
of a run's time must land within 10 % of that run's measured loop time, on
the headline full electrode (466 M voxels, FP32, ~8 s) and on the standard
tier of the CLI (34^2 x 210, a launch-bound run of a few tens of ms). A
budget that covers the whole run gives that run's own time (``exact``).
"""

import pytest

from configs import aic_config
from pulsed_ion_chamber_hip import run_simulation_hip
from pulsed_ion_chamber_hip.benchmark import estimate_full_runtime_empirical

pytestmark = pytest.mark.gpu


def _loop_time(cfg, precision):
    # a second run, so that first-launch costs are not in the figure
    run_simulation_hip(cfg, progress=False, precision=precision, return_fields=False)
    res = run_simulation_hip(cfg, progress=False, precision=precision, max_steps=cfg.total_time_steps,
                             return_fields=False)
    return res.loop_elapsed_s


@pytest.mark.parametrize("where,precision", [((0.265, 5.0), "fp32"), ((0.014, 10.0), "fp64")],
                         ids=["full_electrode_5um_fp32", "standard_tier_fp64"])
def test_estimate_from_a_fifth_of_the_run_is_within_10_percent(where, precision):
    cfg = aic_config(*where)
    measured = _loop_time(cfg, precision)
    est = estimate_full_runtime_empirical(cfg, max_wall_s=0.2 * measured, precision=precision)
    err = abs(est["estimated_seconds"] - measured) / measured
    print(f"{where} {precision}: measured {measured:.4f} s, estimated {est['estimated_seconds']:.4f} s "
          f"from {est['steps_measured']} of {est['total_time_steps']} steps ({err:.1%})")
    assert not est["exact"]
    assert est["steps_measured"] < est["total_time_steps"]
    assert err <= 0.10


def test_budget_covering_the_run_is_exact():
    cfg = aic_config(0.003, 10.0)  # the dev tier, 12^2 x 210
    est = estimate_full_runtime_empirical(cfg, max_wall_s=30.0)
    assert est["exact"]
    assert est["steps_measured"] == est["total_time_steps"] == cfg.total_time_steps
    assert est["estimated_seconds"] == est["elapsed_measured_s"]
