"""Configurations shared by the tests, named after the reference tests that use them."""

from pulsed_ion_chamber_hip.config import SimulationConfig
from pulsed_ion_chamber_hip.constants import (
    AIR_DENSITY_20C_KG_M3,
    ION_DIFFUSION_NEGATIVE_CM2_S,
    ION_DIFFUSION_POSITIVE_CM2_S,
    ION_MOBILITY_NEGATIVE_CM2_VS,
    ION_MOBILITY_POSITIVE_CM2_VS,
)

# tests/test_backends_agree.py:112-126, tests/test_cuda_backend.py:219-233
def fast_config(seed=3, **overrides):
    return SimulationConfig(
        E_MeV_u=1.0, voltage_V=10.0, electrode_gap_cm=0.02, pulse_duration_s=1e-6, dose_rate_Gy_s=1e-8,
        grid_size_um=10.0, sampled_radius_cm=0.01, buffer_radius=5, no_z_electrode=4, seed=seed, **overrides,
    )


# tests/test_v2_physics.py:27-45, tests/test_charge_conservation.py:35-47
SMALL = dict(
    E_MeV_u=56.2, voltage_V=300.0, electrode_gap_cm=0.2, pulse_duration_s=2e-6, repetition_rate_hz=50.0,
    dose_rate_Gy_s=8.91, grid_size_um=40.0, sampled_radius_cm=0.004, buffer_radius=3, no_z_electrode=3, seed=7,
)
TWO_SPECIES = dict(
    mu_positive_cm2_Vs=ION_MOBILITY_POSITIVE_CM2_VS, mu_negative_cm2_Vs=ION_MOBILITY_NEGATIVE_CM2_VS,
    D_positive_cm2_s=ION_DIFFUSION_POSITIVE_CM2_S, D_negative_cm2_s=ION_DIFFUSION_NEGATIVE_CM2_S,
)

# examples/ifj_aic144/run_markus_2mm.py:71-113 with profiling/bench_gpu.py:71-89
AIC_KWARGS = dict(
    E_MeV_u=56.2, voltage_V=300.0, electrode_gap_cm=0.2, pulse_duration_s=540e-6, repetition_rate_hz=50.0,
    n_pulses=1, rf_frequency_hz=26.26e6, seed=20260527, W_eV=33.0, air_density_kg_m3=AIR_DENSITY_20C_KG_M3,
    lateral_boundary="reflecting", n_clearance_separation_times=2.6, track_cutoff_sigmas=10.0, **TWO_SPECIES,
)


def aic_config(radius_cm, grid_um, dose_rate_water_Gy_s=10.0, **overrides):
    kw = dict(AIC_KWARGS, max_voxels=2e9, memory_budget_fraction=None)
    kw.update(overrides)
    return SimulationConfig(
        **kw, dose_rate_Gy_s=dose_rate_water_Gy_s * 0.891, grid_size_um=grid_um, sampled_radius_cm=radius_cm,
        buffer_radius=3, no_z_electrode=5,
    )


# The benchmark workloads: the README example (10 um, r 0.008 cm) and the
# mid-size 5 um, r 0.05 cm solve, library-default physics, 8.91 Gy/s, seed 1.
def readme_config(grid_um=10.0, radius_cm=0.008, **overrides):
    kw = dict(E_MeV_u=56.2, voltage_V=300.0, electrode_gap_cm=0.2, pulse_duration_s=540e-6, repetition_rate_hz=50.0,
              dose_rate_Gy_s=8.91, grid_size_um=grid_um, sampled_radius_cm=radius_cm, seed=1,
              memory_budget_fraction=None)
    kw.update(overrides)
    return SimulationConfig(**kw)


# Saturation curve: 8 log-spaced voltages x 4 dose rates, the README beam and
# chamber otherwise; each point gets a grid from the divisors of the 2 mm gap
# and a sampled radius drawn from default_rng(2026), so the 32 grids and step
# counts are uneven.
CURVE_VOLTAGES_V = tuple(round(50.0 * 10.0 ** (k / 7.0), 2) for k in range(8))
CURVE_DOSE_RATES_GY_S = (2.0, 8.91, 20.0, 40.0)
CURVE_GRIDS_UM = (10.0, 12.5, 16.0, 20.0)


def saturation_curve_configs(seed=2026, voltages=CURVE_VOLTAGES_V, dose_rates=CURVE_DOSE_RATES_GY_S):
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for voltage in voltages:
        for dose in dose_rates:
            grid = CURVE_GRIDS_UM[int(rng.integers(len(CURVE_GRIDS_UM)))]
            radius = float(rng.uniform(0.005, 0.02))
            out.append(readme_config(grid, radius, voltage_V=voltage, dose_rate_Gy_s=dose))
    return out
