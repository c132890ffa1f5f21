"""Pins for the CPU oracle, so that the oracle the HIP backend is checked
against is itself checked against the reference's published numbers.

What pins what:
* config restatement (config.py:43-70, 180-401): grid, steps and track count
  of every published run and every tier; dt against the archived CSV.
* schedule + sampler + deposition + sweep + scoring + wall (the whole loop):
  the C oracle reproduces the reference's published full-precision k_s and
  200-step f values bit for bit (profiling/data/gpu_athena, profiling/data/gh200).
* sampler: the blocked sampler equals the one-attempt-at-a-time definition
  across uneven batches (tests/test_grid_and_timing.py:98-121 of the reference).
* serial NumPy oracle vs C batched oracle: the reference's cross-backend
  tolerances (tests/test_backends_agree.py), here much tighter.
* Jaffe theory: the reference's golden value (web/core/src/theory.rs:76-82).
* planted mistakes in the oracle (swapped drift direction, dropped
  recombination, scoring one gap layer short) each move k_s far outside the
  published value, so the pins above are sensitive to exactly such errors.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES, aic_config, fast_config
from oracle import batched, serial
from oracle.theory import jaffe_ks
from pulsed_ion_chamber_hip.config import SimulationConfig
from pulsed_ion_chamber_hip.pulses import CylinderSampler, draw_tracks, sample_xy_inside_cylinder
from pulsed_ion_chamber_hip.state import build_result, track_density_xy
from pulsed_ion_chamber_hip.stopping_power import E_MeV_u_to_LET_keV_um

PUBLISHED = json.loads((Path(__file__).parent / "golden" / "published.json").read_text())


@pytest.mark.parametrize("row", PUBLISHED["aic_runs"], ids=lambda r: f"r{r['sampled_radius_cm']}@{r['grid_size_um']}")
def test_config_reproduces_every_published_grid(row):
    c = aic_config(row["sampled_radius_cm"], row["grid_size_um"])
    assert c.no_xy == row["no_xy"]
    assert c.no_z_with_buffer == row["no_z_with_buffer"]
    assert c.total_time_steps == row["total_time_steps"]
    if "tracks_per_pulse" in row:
        assert c.number_of_tracks_per_pulse == row["tracks_per_pulse"]


@pytest.mark.parametrize("tier", PUBLISHED["tiers"], ids=lambda t: t["tier"])
def test_config_reproduces_the_tier_table(tier):
    c = aic_config(tier["sampled_radius_cm"], 10.0)
    assert (c.no_xy, c.no_z_with_buffer, c.number_of_tracks_per_pulse) == (tier["no_xy"], tier["nzb"], tier["tracks"])


def test_dt_matches_the_archived_record():
    c = aic_config(0.265, 10.0)
    # the CSV prints time = (step+1)*dt with 13 significant digits
    assert c.dt == pytest.approx(PUBLISHED["full_electrode_10um_first_time_s"], rel=1e-12)
    assert c.total_time_steps == PUBLISHED["full_electrode_10um_steps"]


def _published_ks(radius, grid):
    for row in PUBLISHED["aic_runs"]:
        if row["sampled_radius_cm"] == radius and row["grid_size_um"] == grid and "cpu_ks" in row:
            return row["cpu_ks"]
    raise KeyError((radius, grid))


def test_oracle_reproduces_published_ks_standard_tier_bit_for_bit():
    res = batched.run_batched(aic_config(0.014, 10.0), fields=False)
    assert res.ks == _published_ks(0.014, 10.0)


def test_oracle_reproduces_published_200_step_prefix_bit_for_bit():
    row = next(r for r in PUBLISHED["prefix_f"] if r["grid_size_um"] == 10.0)
    res = batched.run_batched(aic_config(0.05, 10.0), n_steps=row["steps"], fields=False)
    assert res.f_t[row["steps"] - 1] == row["f_end"]
    assert np.isnan(res.ks)


@pytest.mark.slow
def test_oracle_reproduces_published_ks_r005_bit_for_bit():
    res = batched.run_batched(aic_config(0.05, 10.0), fields=False)
    assert res.ks == _published_ks(0.05, 10.0)


def _run_with(config, **param_overrides):
    schedule, xs, ys = draw_tracks(config)
    q = batched.params(config)
    for key, value in param_overrides.items():
        setattr(q, key, value)
    real = batched.params
    batched.params = lambda cfg, alpha_cm3_s=1.60e-6: q
    try:
        per_step, _, _ = batched.run_per_step(config, schedule, xs, ys, fields=False)
    finally:
        batched.params = real
    return build_result(config, per_step, config.total_time_steps, None, None,
                        track_density_xy(config, xs, ys), truncated=False).ks


@pytest.mark.parametrize(
    "mistake",
    ["swap_positive_drift", "no_recombination", "score_from_second_gap_layer", "reflect_as_absorb"],
)
def test_planted_oracle_mistakes_miss_the_published_ks(mistake):
    c = aic_config(0.014, 10.0)
    q = batched.params(c)
    overrides = {
        "swap_positive_drift": dict(p_zm=q.p_zp, p_zp=q.p_zm),
        "no_recombination": dict(alpha_dt=0.0),
        "score_from_second_gap_layer": dict(nze=q.nze + 1, no_z=q.no_z - 1),
        "reflect_as_absorb": dict(reflecting=0),
    }[mistake]
    ks = _run_with(c, **overrides)
    assert abs(ks - _published_ks(0.014, 10.0)) > 1e-6


def test_jaffe_matches_the_reference_golden():
    ks = jaffe_ks(E_MeV_u_to_LET_keV_um(1.0), 10.0, 0.02)
    assert ks == pytest.approx(PUBLISHED["jaffe_ks_single_track_golden"], rel=0, abs=5e-15)
    assert ks == pytest.approx(PUBLISHED["jaffe_ks_validation_json"], rel=0, abs=5e-15)


def test_jaffe_honours_a_non_default_w_ev():
    default_loss = jaffe_ks(60.0, 300.0, 0.2) - 1.0
    overridden_loss = jaffe_ks(60.0, 300.0, 0.2, W_eV=33.0) - 1.0
    assert overridden_loss / default_loss == pytest.approx(34.2 / 33.0, rel=0.05)


def test_block_sampler_draws_exactly_what_the_per_track_sampler_would():
    mid_xy, radius, no_xy = 30, 24.0, 60
    batches = [7, 1, 5000, 3, 900]
    reference = np.random.default_rng(11)
    expected = np.array([sample_xy_inside_cylinder(reference, mid_xy, radius, no_xy) for _ in range(sum(batches))])
    sampler = CylinderSampler(np.random.default_rng(11), mid_xy, radius, no_xy)
    got = np.vstack([np.column_stack(sampler.sample(n)) for n in batches])
    np.testing.assert_array_equal(got, expected)


def test_block_sampler_stays_inside_the_disc():
    xs, ys = CylinderSampler(np.random.default_rng(2), 30, 24.0, 60).sample(20_000)
    assert len(xs) == len(ys) == 20_000
    assert np.all((xs - 30) ** 2 + (ys - 30) ** 2 <= 24.0**2)


def test_whole_run_sampling_equals_step_by_step_sampling():
    c = SimulationConfig(**SMALL)
    schedule, xs, ys = draw_tracks(c)
    rng = np.random.default_rng(c.seed)
    from pulsed_ion_chamber_hip.pulses import build_track_schedule

    sched2 = build_track_schedule(c, rng)
    sampler = CylinderSampler(rng, c.mid_xy, c.sampling_radius, c.no_xy)
    parts = [sampler.sample(n) for n in sched2]
    np.testing.assert_array_equal(schedule, sched2)
    np.testing.assert_array_equal(xs, np.concatenate([p[0] for p in parts]))
    np.testing.assert_array_equal(ys, np.concatenate([p[1] for p in parts]))


@pytest.mark.parametrize(
    "cfg",
    [
        fast_config(),
        fast_config(lateral_boundary="reflecting"),
        SimulationConfig(**SMALL),
        SimulationConfig(**SMALL, **TWO_SPECIES, lateral_boundary="reflecting", scoring_region="full_grid"),
        SimulationConfig(**SMALL, chamber_fill_fraction=0.7),
        SimulationConfig(**SMALL, **TWO_SPECIES, track_cutoff_sigmas=None),
    ],
    ids=["fast", "fast-reflecting", "small", "small-all-options", "small-fill", "small-nocutoff"],
)
def test_serial_and_batched_oracles_agree(cfg):
    s = serial.run_serial(cfg)
    b = batched.run_batched(cfg)
    peak = np.abs(s.positive_array).max()
    assert b.ks == pytest.approx(s.ks, rel=1e-12)
    np.testing.assert_allclose(b.f_t, s.f_t, rtol=1e-12)
    np.testing.assert_allclose(b.positive_array, s.positive_array, rtol=0, atol=1e-12 * peak)
    np.testing.assert_allclose(b.negative_array, s.negative_array, rtol=0, atol=1e-12 * peak)
    np.testing.assert_array_equal(b.track_density_xy, s.track_density_xy)


def test_one_sweep_of_the_two_oracles_is_bit_identical():
    """With identical inputs the serial and C sweeps perform the same
    operations in the same order, so the fields must agree exactly."""
    cfg = SimulationConfig(**SMALL, **TWO_SPECIES)
    rng = np.random.default_rng(5)
    shape = (cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer)
    pos = rng.random(shape) * 1e7
    neg = rng.random(shape) * 1e7
    pn, nn, _ = batched.lw_step(cfg, pos, neg)
    ps, ns = np.zeros(shape), np.zeros(shape)
    serial.lw_step(pos, neg, ps, ns, cfg, 1.60e-6 * cfg.dt)
    np.testing.assert_array_equal(pn[1:-1, 1:-1, 1:-1], ps[1:-1, 1:-1, 1:-1])
    np.testing.assert_array_equal(nn[1:-1, 1:-1, 1:-1], ns[1:-1, 1:-1, 1:-1])
