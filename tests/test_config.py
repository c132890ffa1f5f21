"""The input contract, ported from the reference's config tests
(tests/test_grid_and_timing.py:1-121, tests/test_v2_physics.py:48-240)."""

import warnings

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES
from pulsed_ion_chamber_hip.config import SimulationConfig
from pulsed_ion_chamber_hip.constants import (
    AIR_DENSITY_KG_M3,
    ION_DIFFUSION_CM2_S,
    ION_DIFFUSION_NEGATIVE_CM2_S,
    ION_MOBILITY_CM2_VS,
    ION_MOBILITY_NEGATIVE_CM2_VS,
    ION_MOBILITY_POSITIVE_CM2_VS,
    W_EV_PER_ION_PAIR,
)


def test_default_config_builds_and_is_small():
    config = SimulationConfig(seed=0)
    assert config.no_xy > 2 * config.buffer_radius
    assert config.no_z > 0 and config.dt > 0 and config.number_of_tracks_per_pulse >= 1
    assert config.no_xy * config.no_xy * config.no_z_with_buffer < 1e6


def test_pulse_period_must_fit_pulse_duration():
    with pytest.raises(ValueError):
        SimulationConfig(pulse_duration_s=1e-2, repetition_rate_hz=1000.0, n_pulses=2)


def test_grid_size_guard_raises_for_huge_grid():
    with pytest.raises(ValueError):
        SimulationConfig(grid_size_um=0.1, sampled_radius_cm=1.0, electrode_gap_cm=1.0)


def test_more_pulses_take_more_total_time_steps():
    assert SimulationConfig(n_pulses=2).total_time_steps > SimulationConfig(n_pulses=1).total_time_steps


def test_degenerate_inner_radius_raises_instead_of_hanging():
    with pytest.raises(ValueError, match="inner_radius"):
        SimulationConfig(grid_size_um=100.0, sampled_radius_cm=0.001, buffer_radius=2, no_z_electrode=2)


def test_rf_cycles_per_time_step():
    assert SimulationConfig(seed=0).rf_cycles_per_time_step is None
    config = SimulationConfig(rf_frequency_hz=26.26e6, seed=0)
    assert config.rf_cycles_per_time_step == pytest.approx(config.dt * 26.26e6)
    with pytest.warns(UserWarning, match="RF"):
        SimulationConfig(rf_frequency_hz=1.0, seed=0)


def test_gap_mismatch_warns():
    with pytest.warns(UserWarning, match="electrode_gap_cm"):
        config = SimulationConfig(grid_size_um=40.0, electrode_gap_cm=0.05, seed=0)
    assert config.no_z == 12
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert SimulationConfig(grid_size_um=10.0, electrode_gap_cm=0.7, seed=0).no_z == 700


def test_defaults_are_the_single_averaged_species():
    config = SimulationConfig(**SMALL)
    assert not config.two_carrier_species
    assert config.mu_positive == config.mu_negative == ION_MOBILITY_CM2_VS
    assert config.D_positive == config.D_negative == ION_DIFFUSION_CM2_S
    assert config.W_eV == W_EV_PER_ION_PAIR and config.air_density_kg_m3 == AIR_DENSITY_KG_M3
    pos, neg = config.scheme_coefficients()
    assert pos == neg
    assert config.sampling_radius == config.inner_radius
    assert config.scoring_radius_sq == config.inner_radius_sq


def test_two_species_dt_is_set_by_the_negative_ion():
    single = SimulationConfig(**SMALL)
    both = SimulationConfig(**SMALL, **TWO_SPECIES)
    assert both.two_carrier_species and both.dt < single.dt
    assert both.total_time_steps > single.total_time_steps
    negative_only = SimulationConfig(
        **SMALL, mu_positive_cm2_Vs=ION_MOBILITY_NEGATIVE_CM2_VS, mu_negative_cm2_Vs=ION_MOBILITY_NEGATIVE_CM2_VS,
        D_positive_cm2_s=ION_DIFFUSION_NEGATIVE_CM2_S, D_negative_cm2_s=ION_DIFFUSION_NEGATIVE_CM2_S,
    )
    assert both.dt == negative_only.dt
    assert both.slowest_mobility_cm2_Vs == ION_MOBILITY_POSITIVE_CM2_VS


def test_von_neumann_criterion_holds_for_both_species():
    config = SimulationConfig(**SMALL, **TWO_SPECIES)
    for diffusion, mobility in ((config.D_positive, config.mu_positive), (config.D_negative, config.mu_negative)):
        s = diffusion * config.dt / config.unit_length_cm**2
        c = mobility * config.Efield_V_cm * config.dt / config.unit_length_cm
        assert 6 * s + c**2 <= 1.0


def test_no_xy_rounds_so_the_sampled_radius_is_honoured():
    config = SimulationConfig(**{**SMALL, "grid_size_um": 10.0, "sampled_radius_cm": 0.0084})
    assert config.no_xy == 17 + 2 * config.buffer_radius


def test_chamber_fill_fraction_shrinks_placement_not_scoring():
    full = SimulationConfig(**SMALL)
    partial = SimulationConfig(**SMALL, chamber_fill_fraction=0.7)
    assert partial.number_of_tracks_per_pulse == full.number_of_tracks_per_pulse
    assert partial.scoring_radius_sq == full.scoring_radius_sq
    assert partial.sampling_radius == pytest.approx(0.7 * full.inner_radius)


@pytest.mark.parametrize("fraction", [0.0, -0.1, 1.5])
def test_invalid_chamber_fill_fraction_rejected(fraction):
    with pytest.raises(ValueError, match="chamber_fill_fraction"):
        SimulationConfig(**SMALL, chamber_fill_fraction=fraction)


@pytest.mark.parametrize(
    "kwargs, match",
    [
        (dict(lateral_boundary="sticky"), "lateral_boundary"),
        (dict(scoring_region="everything"), "scoring_region"),
        (dict(mu_negative_cm2_Vs=-2.1), "mu_negative_cm2_Vs"),
        (dict(track_cutoff_sigmas=0.0), "track_cutoff_sigmas"),
        (dict(dt_seconds=1.0), "dt_seconds"),
    ],
)
def test_invalid_options_rejected(kwargs, match):
    with pytest.raises(ValueError, match=match):
        SimulationConfig(**SMALL, **kwargs)


def test_full_grid_scoring_admits_every_voxel():
    config = SimulationConfig(**SMALL, scoring_region="full_grid")
    corner_sq = 2 * max(config.mid_xy, config.no_xy - 1 - config.mid_xy) ** 2
    assert config.scoring_radius_sq > corner_sq


def test_cutoff_geometry_and_discarded_charge():
    config = SimulationConfig(**SMALL, track_cutoff_sigmas=10.0)
    assert config.track_sigma_cm == pytest.approx(config.track_radius_cm / np.sqrt(2.0))
    assert config.track_cutoff_voxels == pytest.approx(10.0 * config.track_sigma_cm / config.unit_length_cm)
    assert config.truncated_charge_fraction == pytest.approx(np.exp(-50.0))
    none = SimulationConfig(**SMALL, track_cutoff_sigmas=None)
    assert none.truncated_charge_fraction == 0.0 and none.track_cutoff_voxels >= none.no_xy


def test_memory_estimate_matches_the_arrays_that_get_allocated():
    config = SimulationConfig(**SMALL)
    voxels = config.no_xy**2 * config.no_z_with_buffer
    assert config.carrier_array_bytes == 4 * voxels * 8
    assert config.track_schedule_bytes == 2 * config.number_of_tracks_per_pulse * 8
    assert config.estimated_memory_bytes == max(config.track_schedule_bytes,
                                                config.carrier_array_bytes + config.scratch_bytes)


def test_oversized_grid_is_refused_before_it_can_be_allocated():
    with pytest.raises(MemoryError, match="available"):
        SimulationConfig(**{**SMALL, "grid_size_um": 10.0, "sampled_radius_cm": 20.0}, max_voxels=1e14)


def test_memory_check_can_be_disabled():
    config = SimulationConfig(**{**SMALL, "grid_size_um": 10.0, "sampled_radius_cm": 2.0}, max_voxels=1e14,
                              memory_budget_fraction=None)
    assert config.estimated_memory_bytes > 0
