"""Per-step diagnostics of a HIP run: the reference's checks on the Result
series (tests/test_output_and_diagnostics.py:18-100) and on the track cutoff
(tests/test_v2_physics.py:214-219), run on run_simulation_hip with the same
configuration and tolerances."""

import numpy as np
import pytest

from pulsed_ion_chamber_hip import SimulationConfig, run_simulation_hip
from pulsed_ion_chamber_hip.output import track_density_per_cm2

pytestmark = pytest.mark.gpu

# tests/test_output_and_diagnostics.py:18-30
SMALL = dict(
    E_MeV_u=56.2,
    voltage_V=300.0,
    electrode_gap_cm=0.2,
    pulse_duration_s=4e-6,
    repetition_rate_hz=50.0,
    dose_rate_Gy_s=8.91,
    grid_size_um=40.0,
    sampled_radius_cm=0.004,
    buffer_radius=3,
    no_z_electrode=3,
    seed=11,
)


@pytest.fixture(scope="module")
def result():
    return run_simulation_hip(SimulationConfig(**SMALL), progress=False)


def test_diagnostics_reproduce_the_collection_efficiency(result):
    injected = result.injected_positive.cumsum()
    recombined = result.recombination.cumsum()
    expected = (injected - recombined) / injected
    np.testing.assert_allclose(result.f_t[injected > 0], expected[injected > 0], rtol=1e-12)
    assert result.ks == pytest.approx(1.0 / expected[-1], rel=1e-12)


def test_both_signs_are_injected_equally(result):
    np.testing.assert_array_equal(result.injected_positive, result.injected_negative)


def test_injection_is_confined_to_the_pulse(result):
    pulse_steps = result.config.pulse_time_steps
    assert result.injected_positive[:pulse_steps].sum() > 0
    assert result.injected_positive[pulse_steps:].sum() == 0


def test_carriers_are_collected_by_the_end_of_the_run(result):
    peak = result.n_positive.max()
    assert abs(result.n_positive[-1]) < 1e-3 * peak
    assert abs(result.n_negative[-1]) < 1e-3 * peak


def test_extra_clearance_empties_the_gap_without_moving_k_s(result):
    with_margin = run_simulation_hip(SimulationConfig(**SMALL, n_clearance_separation_times=2.6), progress=False)
    peak = result.n_positive.max()
    assert abs(result.n_positive[-1]) > 1e-9 * peak  # no margin: something is left
    assert abs(with_margin.n_positive[-1]) < 1e-9 * peak  # margin: nothing is
    assert with_margin.ks == pytest.approx(result.ks, rel=1e-6)


def test_all_track_centres_are_counted_inside_the_sampling_disc(result):
    config = result.config
    assert result.track_density_xy.sum() == config.number_of_tracks_per_pulse
    i, j = np.nonzero(result.track_density_xy)
    radius = np.hypot(i - config.mid_xy, j - config.mid_xy)
    assert radius.max() <= config.sampling_radius + np.sqrt(2)


def test_track_areal_density_matches_the_nominal_fluence(result):
    config = result.config
    total = track_density_per_cm2(result).sum() * config.unit_length_cm**2
    assert total == pytest.approx(config.number_of_tracks_per_pulse)


def test_default_cutoff_is_exact_to_machine_precision():
    """10 sigma discards 1.9e-22 of each track: it must not move k_s relative
    to depositing over the whole grid (tests/test_v2_physics.py:214-219)."""
    full = run_simulation_hip(SimulationConfig(**SMALL, track_cutoff_sigmas=None), progress=False)
    stencil = run_simulation_hip(SimulationConfig(**SMALL, track_cutoff_sigmas=10.0), progress=False)
    assert stencil.ks == pytest.approx(full.ks, rel=1e-12)
