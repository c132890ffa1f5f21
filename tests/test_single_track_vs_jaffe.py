"""Single-track limit against Jaffe theory (the reference's
tests/test_single_track_vs_jaffe.py:17-86), on the HIP backend; the oracle runs
the same check on the CPU."""

import numpy as np
import pytest

from oracle import batched
from oracle.theory import jaffe_ks
from pulsed_ion_chamber_hip import SimulationConfig, run_simulation_hip


def _config(seed):
    return SimulationConfig(
        E_MeV_u=1.0, voltage_V=10.0, electrode_gap_cm=0.02, pulse_duration_s=1e-6, dose_rate_Gy_s=1e-8,
        n_pulses=1, grid_size_um=10.0, sampled_radius_cm=0.01, buffer_radius=5, no_z_electrode=4, seed=seed,
    )


def _jaffe(config):
    return jaffe_ks(config.LET_keV_um, config.voltage_V, config.electrode_gap_cm, W_eV=config.W_eV,
                    mu=config.mu_positive, D=config.D_positive)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_single_track_matches_jaffe_theory(seed):
    config = _config(seed)
    assert config.number_of_tracks_per_pulse == 1
    result = run_simulation_hip(config, progress=False)
    assert (result.ks - 1.0) / (_jaffe(config) - 1.0) == pytest.approx(1.0, rel=0.3)


@pytest.mark.gpu
def test_f_t_is_monotonically_non_increasing():
    result = run_simulation_hip(_config(1), progress=False)
    assert np.all(np.diff(result.f_t) <= 1e-6)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_single_track_matches_jaffe_theory(seed):
    config = _config(seed)
    result = batched.run_batched(config)
    assert (result.ks - 1.0) / (_jaffe(config) - 1.0) == pytest.approx(1.0, rel=0.3)
