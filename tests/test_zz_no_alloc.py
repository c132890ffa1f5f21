"""pic_run_pulse_static, the entry point that allocates nothing, against
pic_run_pulse, the entry point it varies: the same per-step record and the
same fields, bit for bit, on two calls in a row (the second with a smaller
input, so the static scratch holds the first call's larger contents when the
second starts), in FP64 and FP32, through the C ABI. Inputs beyond the
static scratch are refused with a clear error, before anything runs."""

import ctypes

import numpy as np
import pytest

from configs import readme_config
from pulsed_ion_chamber_hip import _lib
from pulsed_ion_chamber_hip.pulses import draw_tracks
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu


def _calls():
    """A 48^2 x 206 grid at 40 Gy/s (~100 tracks per in-pulse step): the whole
    run first, then a shorter, sparser schedule (every other step's tracks
    dropped, 400 steps)."""
    cfg = readme_config(10.0, 0.02, dose_rate_Gy_s=40.0)
    schedule, xs, ys = draw_tracks(cfg)
    first = (schedule, xs, ys)
    offsets = np.concatenate([[0], np.cumsum(schedule)])
    n2 = 400
    keep = np.zeros(len(xs), dtype=bool)
    sched2 = np.zeros(n2, dtype=np.int64)
    for s in range(0, n2, 2):
        keep[offsets[s]:offsets[s + 1]] = True
        sched2[s] = schedule[s]
    second = (sched2, xs[keep], ys[keep])
    assert len(second[1]) < len(xs) and n2 < len(schedule)
    return cfg, first, second


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_static_entry_point_matches_run_pulse_on_two_calls(precision):
    cfg, first, second = _calls()
    with HipSimulation(cfg, precision) as ref, HipSimulation(cfg, precision) as got:
        for schedule, xs, ys in (first, second):
            want = ref.run_pulse(schedule, xs, ys)
            have = got.run_pulse_static(schedule, xs, ys)
            assert have.shape == (len(schedule), 4)
            np.testing.assert_array_equal(have, want)
            assert have[:, 0].sum() > 0
            for a, b in zip(got.read_fields(), ref.read_fields()):
                np.testing.assert_array_equal(a, b)


def test_static_entry_point_leaves_the_uploaded_tracks():
    cfg, first, second = _calls()
    with HipSimulation(cfg, "fp64") as sim:
        sim.upload_tracks(*first)
        sim.run_pulse_static(*second)
        sim.reset()
        sim.run_steps(0, cfg.total_time_steps)
        after = sim.read_per_step()
    with HipSimulation(cfg, "fp64") as sim:
        plain = sim.run_pulse(*first)
    np.testing.assert_array_equal(after, plain)


def test_inputs_beyond_the_static_scratch_are_refused():
    cfg, first, _ = _calls()
    n_tracks = 5_000_000  # 80 MB of centres
    schedule = np.zeros(10, dtype=np.int64)
    schedule[0] = n_tracks
    xs = np.full(n_tracks, cfg.mid_xy)
    with HipSimulation(cfg, "fp64") as sim:
        with pytest.raises(ValueError, match="static scratch"):
            sim.run_pulse_static(schedule, xs, xs)
        # the handle still works, and the refusal changed nothing
        have = sim.run_pulse_static(*first)
    with HipSimulation(cfg, "fp64") as sim:
        want = sim.run_pulse(*first)
    np.testing.assert_array_equal(have, want)
