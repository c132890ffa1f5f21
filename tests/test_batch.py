"""A batched parameter sweep must give every configuration the numbers it
gets when run alone.

The reference runs a sweep as one process per configuration
(docs/PERFORMANCE.md:211-218), so its answer for each point is simply that
configuration's run; the batch is checked against (1) the same configurations
run one at a time through the single-configuration entry point and (2) the
CPU oracle. Grids, step counts and orderings are deliberately uneven.

Fields are compared bit for bit: the voxel update does not depend on the
tiling or on which launch carries it. The per-step scored sums are folded from
per-block partials, whose grouping follows the tiling, so f_t and k_s are
compared at 1e-12 relative.
"""

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES, readme_config
from oracle import batched
from pulsed_ion_chamber_hip import SimulationConfig, run_batch_hip, run_simulation_hip
from pulsed_ion_chamber_hip.batch import HipBatch
from pulsed_ion_chamber_hip.pulses import draw_tracks
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def uneven_configs():
    """Four small saturation-curve points: different grids (one with an odd
    nzb), voltages and so step counts, listed shortest-run first so the
    batch must reorder them."""
    return [
        readme_config(20.0, 0.0146, voltage_V=50.0, dose_rate_Gy_s=2.0),      # 23^2 x 106, 243 steps
        readme_config(16.0, 0.0077, voltage_V=69.47, dose_rate_Gy_s=2.0),     # 18^2 x 131, 360 steps
        readme_config(20.0, 0.0101, voltage_V=96.53, dose_rate_Gy_s=40.0),    # 18^2 x 106, 337 steps
        readme_config(12.5, 0.0052, voltage_V=134.13, dose_rate_Gy_s=40.0),   # 16^2 x 166, 696 steps
    ]


def _check_same(batch_res, single_res):
    np.testing.assert_array_equal(batch_res.positive_array, single_res.positive_array)
    np.testing.assert_array_equal(batch_res.negative_array, single_res.negative_array)
    np.testing.assert_array_equal(batch_res.track_density_xy, single_res.track_density_xy)
    np.testing.assert_allclose(batch_res.f_t, single_res.f_t, rtol=RTOL)
    np.testing.assert_allclose(batch_res.recombination, single_res.recombination, rtol=RTOL, atol=1e-300)
    np.testing.assert_array_equal(batch_res.injected_positive, single_res.injected_positive)
    assert batch_res.ks == pytest.approx(single_res.ks, rel=RTOL)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_matches_each_configuration_run_alone(precision):
    cfgs = uneven_configs()
    assert len({c.total_time_steps for c in cfgs}) == len(cfgs)
    got = run_batch_hip(cfgs, precision=precision, return_fields=True)
    for cfg, res in zip(cfgs, got):
        alone = run_simulation_hip(cfg, progress=False, precision=precision)
        _check_same(res, alone)


def _field_tolerance(cfg, pos, neg):
    """1e-12 of the peak, or twice the field change the oracle itself shows
    when every track centre moves by one ulp, whichever is larger.

    The deposition factors come from exp(), which neither the device library
    nor glibc rounds correctly, so the two sides differ by up to an ulp per
    factor. Most grids damp that to ~1e-15 of the peak; at high dose rate and
    low voltage (96.53 V, 40 Gy/s, 20 um) charge piles up against the
    collecting electrode and the same ulp reaches ~5e-11 of the peak in the
    oracle alone, while the scored per-step sums stay at ~2e-15."""
    schedule, xs, ys = draw_tracks(cfg)
    _, p2, n2 = batched.run_per_step(cfg, schedule, np.nextafter(xs, np.inf), ys)
    peak = np.abs(pos).max()
    sens = max(float(np.abs(p2 - pos).max()), float(np.abs(n2 - neg).max()))
    return max(RTOL * peak, 2.0 * sens)


def test_batch_matches_the_oracle_in_fp64():
    cfgs = uneven_configs()
    got = run_batch_hip(cfgs, precision="fp64", return_fields=True)
    for cfg, res in zip(cfgs, got):
        ref = batched.run_batched(cfg)
        tol = _field_tolerance(cfg, ref.positive_array, ref.negative_array)
        assert float(np.abs(res.positive_array - ref.positive_array).max()) <= tol
        assert float(np.abs(res.negative_array - ref.negative_array).max()) <= tol
        np.testing.assert_allclose(res.f_t, ref.f_t, rtol=RTOL)
        assert res.ks == pytest.approx(ref.ks, rel=RTOL)


def test_fp32_batch_keeps_ks_within_1e6_of_the_oracle():
    cfgs = uneven_configs()
    got = run_batch_hip(cfgs, precision="fp32")
    for cfg, res in zip(cfgs, got):
        ref = batched.run_batched(cfg, fields=False)
        assert abs(res.ks - ref.ks) <= 1e-6


def test_reflecting_two_species_batch_matches_the_oracle():
    cfgs = [
        SimulationConfig(**{**SMALL, **TWO_SPECIES, "lateral_boundary": "reflecting"}),
        SimulationConfig(**{**SMALL, **TWO_SPECIES, "lateral_boundary": "reflecting", "voltage_V": 150.0,
                            "seed": 11}),
        SimulationConfig(**{**SMALL, **TWO_SPECIES, "lateral_boundary": "reflecting", "grid_size_um": 25.0}),
    ]
    got = run_batch_hip(cfgs, precision="fp64", return_fields=True)
    for cfg, res in zip(cfgs, got):
        ref = batched.run_batched(cfg)
        peak = np.abs(ref.positive_array).max()
        assert float(np.abs(res.positive_array - ref.positive_array).max()) <= RTOL * peak
        assert float(np.abs(res.negative_array - ref.negative_array).max()) <= RTOL * peak
        assert res.ks == pytest.approx(ref.ks, rel=RTOL)


@pytest.mark.parametrize("tiling", [(1, 4), (2, 8), (4, 16), (1, 2), (1, 1)])
def test_every_batch_tiling_gives_the_same_fields(tiling):
    """(1, 2) and (1, 1) run the short march (exactly one or two planes per
    block), the others the general one."""
    cfgs = uneven_configs()[:3]
    base = run_batch_hip(cfgs, precision="fp32", return_fields=True)
    got = run_batch_hip(cfgs, precision="fp32", return_fields=True, tiling=tiling)
    for a, b in zip(got, base):
        np.testing.assert_array_equal(a.positive_array, b.positive_array)
        np.testing.assert_array_equal(a.negative_array, b.negative_array)
        assert a.ks == pytest.approx(b.ks, rel=RTOL)


def three_lane_configs():
    """Eight small curve points whose run lengths split a batch into three
    lanes: one run over 0.8 of the longest (1133 steps), three over 0.36 of
    it (432-696), four shorter (189-337); rows of one and two k windows."""
    pts = [(12.5, 0.0077, 258.97), (12.5, 0.0101, 134.13), (16.0, 0.0077, 134.13), (16.0, 0.0052, 96.53),
           (20.0, 0.0101, 96.53), (16.0, 0.0077, 50.0), (20.0, 0.0077, 50.0), (25.0, 0.0052, 50.0)]
    return [readme_config(g, r, voltage_V=v, dose_rate_Gy_s=8.91) for g, r, v in pts]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_a_batch_in_three_lanes_matches_each_configuration_run_alone(precision):
    cfgs = three_lane_configs()
    steps = np.array([c.total_time_steps for c in cfgs])
    ratio = steps / steps.max()
    # the library's split: over 0.8 of the longest, over 0.36, the rest
    assert (ratio > 0.8).sum() == 1 and ((ratio > 0.36) & (ratio <= 0.8)).sum() == 3 and (ratio <= 0.36).sum() == 4
    got = run_batch_hip(cfgs, precision=precision, return_fields=True)
    for cfg, res in zip(cfgs, got):
        _check_same(res, run_simulation_hip(cfg, progress=False, precision=precision))


def test_a_batch_run_in_pieces_equals_one_run():
    cfgs = uneven_configs()
    draws = [draw_tracks(c) for c in cfgs]

    def run(pieces):
        sims = [HipSimulation(c, "fp64") for c in cfgs]
        for sim, d in zip(sims, draws):
            sim.upload_tracks(*d)
        with HipBatch(sims) as b:
            for first, count in pieces:
                b.run(first, count)
            b.synchronize()
        out = [(s.read_per_step(), *s.read_fields()) for s in sims]
        for s in sims:
            s.close()
        return out

    longest = max(c.total_time_steps for c in cfgs)
    whole = run([(0, longest)])
    split = run([(0, 100), (100, 150), (250, longest - 250)])  # 250 runs past the shortest (243)
    for (ra, pa, na), (rb, pb, nb) in zip(whole, split):
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(na, nb)
        np.testing.assert_allclose(ra, rb, rtol=RTOL, atol=1e-300)


def test_batch_rejects_mixed_precision_and_walls():
    a = SimulationConfig(**SMALL)
    b = SimulationConfig(**{**SMALL, "lateral_boundary": "reflecting"})
    with HipSimulation(a, "fp64") as s1, HipSimulation(a, "fp32") as s2, HipSimulation(b, "fp64") as s3:
        with pytest.raises(ValueError, match="precision"):
            HipBatch([s1, s2])
        with pytest.raises(ValueError, match="wall"):
            HipBatch([s1, s3])
        with pytest.raises(ValueError, match="twice"):
            HipBatch([s1, s1])
    with pytest.raises(ValueError, match="lateral boundary"):
        run_batch_hip([a, b])
