"""Kernel-level conservation and parity of the HIP backend.

Port of the reference's tests/test_charge_conservation.py:51-180: deposit 30
tracks and sweep once through the device kernels directly, then check each
kernel's own scored total against an independent NumPy reduction over the
array it just wrote, on a voxel set built here from scratch. Added: the same
deposit and sweep compared with the oracle, where the FP64 sweep must be
bit-identical for identical input, and the lateral ring after a sweep under
both walls.
"""

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES
from oracle import batched, serial
from pulsed_ion_chamber_hip.config import SCORING_REGIONS, SimulationConfig
from pulsed_ion_chamber_hip.pulses import CylinderSampler
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu
N_TRACKS = 30
CC = {**SMALL, "seed": 11}


def _scored_mask(config):
    i = np.arange(config.no_xy)[:, None]
    j = np.arange(config.no_xy)[None, :]
    disc = (i - config.mid_xy) ** 2 + (j - config.mid_xy) ** 2 < config.scoring_radius_sq
    interior = np.zeros_like(disc)
    interior[1:-1, 1:-1] = True
    return disc & interior


def _tracks(config):
    rng = np.random.default_rng(config.seed)
    return CylinderSampler(rng, config.mid_xy, config.sampling_radius, config.no_xy).sample(N_TRACKS)


def _deposit_and_sweep(config, precision="fp64"):
    xs, ys = _tracks(config)
    with HipSimulation(config, precision) as sim:
        inserted = sim.deposit(xs, ys)
        pos, neg = sim.read_fields()
        rec, tp, tn = sim.step()
        pos_next, neg_next = sim.read_fields()
    return dict(positive=pos, negative=neg, positive_next=pos_next, negative_next=neg_next,
                alpha_dt=1.60e-6 * config.dt, inserted=inserted, recombined=rec, total_positive=tp,
                total_negative=tn, xs=xs, ys=ys)


@pytest.mark.parametrize("scoring_region", SCORING_REGIONS)
def test_injected_charge_matches_an_independent_reduction_over_the_scored_voxels(scoring_region):
    config = SimulationConfig(**CC, scoring_region=scoring_region)
    run = _deposit_and_sweep(config)
    gap = slice(config.no_z_electrode, config.no_z_electrode + config.no_z)
    expected = run["positive"][:, :, gap][_scored_mask(config)].sum()
    assert run["inserted"] == pytest.approx(expected, rel=1e-12)


@pytest.mark.parametrize("scoring_region", SCORING_REGIONS)
def test_recombination_and_present_charge_are_scored_over_the_same_gap_layers_as_injection(scoring_region):
    config = SimulationConfig(**CC, scoring_region=scoring_region)
    run = _deposit_and_sweep(config)
    mask = _scored_mask(config)
    gap = slice(config.no_z_electrode, config.no_z_electrode + config.no_z)
    expected_rec = (run["alpha_dt"] * run["positive"] * run["negative"])[:, :, gap][mask].sum()
    assert run["recombined"] == pytest.approx(expected_rec, rel=1e-12)
    assert run["total_positive"] == pytest.approx(run["positive_next"][:, :, gap][mask].sum(), rel=1e-12)
    assert run["total_negative"] == pytest.approx(run["negative_next"][:, :, gap][mask].sum(), rel=1e-12)


def test_deposited_field_matches_the_serial_oracle_track_by_track():
    config = SimulationConfig(**CC, **TWO_SPECIES)
    run = _deposit_and_sweep(config)
    shape = run["positive"].shape
    pos, neg = np.zeros(shape), np.zeros(shape)
    inserted = sum(serial.deposit_track(pos, neg, x, y, config) for x, y in zip(run["xs"], run["ys"]))
    peak = pos.max()
    np.testing.assert_allclose(run["positive"], pos, rtol=0, atol=1e-13 * peak)
    np.testing.assert_allclose(run["negative"], neg, rtol=0, atol=1e-13 * peak)
    assert run["inserted"] == pytest.approx(inserted, rel=1e-13)


def test_deposit_matches_the_batched_oracle_and_is_deterministic():
    config = SimulationConfig(**CC, **TWO_SPECIES, track_cutoff_sigmas=None)
    xs, ys = _tracks(config)
    shape = (config.no_xy, config.no_xy, config.no_z_with_buffer)
    pos, neg = np.zeros(shape), np.zeros(shape)
    ref = batched.deposit(config, pos, neg, xs, ys)
    outs = []
    for _ in range(2):
        with HipSimulation(config, "fp64") as sim:
            got = sim.deposit(xs, ys)
            outs.append((got, *sim.read_fields()))
    assert outs[0][0] == outs[1][0]
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    np.testing.assert_allclose(outs[0][1], pos, rtol=0, atol=1e-14 * pos.max())
    assert outs[0][0] == pytest.approx(ref, rel=1e-13)


# 40 um: 8 x 8 x 56 voxels; 16 um: 11 x 11 x 131, an odd number of k voxels
# (the device pads FP32 rows to an even length)
GRIDS_UM = [40.0, 16.0]

# FP32 k-window layouts, as (grid um, electrode gap cm): one window, an odd
# row length, 16 x 16 x 254 (252 interior k values: two-value windows step by
# 126, two of them instead of three of 124, with single-value edge stores),
# 16 x 16 x 256 (three windows of 124; two four-value windows), 22 x 22 x 406
# (four windows of 124) and 26 x 26 x 504 (four of 126; three four-value
# windows)
# row lengths (nzb) 56, 131, 254, 253, 256, 406, 504: every residue mod 4
FP32_GRIDS = [(40.0, 0.2), (16.0, 0.2), (8.0, 0.1984), (8.0, 0.1976), (8.0, 0.2), (5.0, 0.2), (4.0, 0.1992)]


@pytest.mark.parametrize("grid_um", GRIDS_UM)
@pytest.mark.parametrize("wall", ["absorbing", "reflecting"])
@pytest.mark.parametrize("species", ["one", "two"])
def test_one_fp64_sweep_is_bit_identical_to_the_oracle(wall, species, grid_um):
    """Same input, same operations in the same order, no fused multiply-adds:
    the device sweep must reproduce the oracle's interior exactly."""
    extra = TWO_SPECIES if species == "two" else {}
    config = SimulationConfig(**{**CC, "grid_size_um": grid_um}, **extra, lateral_boundary=wall)
    rng = np.random.default_rng(17)
    shape = (config.no_xy, config.no_xy, config.no_z_with_buffer)
    pos = rng.random(shape) * 1e7
    neg = rng.random(shape) * 1e7
    pos[:, :, [0, -1]] = 0.0
    neg[:, :, [0, -1]] = 0.0
    # every reachable state under a reflecting wall has a mirrored ring
    serial.apply_lateral_boundary(pos, wall)
    serial.apply_lateral_boundary(neg, wall)
    ref_p, ref_n, ref3 = batched.lw_step(config, pos, neg)
    with HipSimulation(config, "fp64") as sim:
        sim.write_fields(pos, neg)
        got3 = sim.step()
        gp, gn = sim.read_fields()
    c = (slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_array_equal(gp[c], ref_p[c])
    np.testing.assert_array_equal(gn[c], ref_n[c])
    np.testing.assert_allclose(got3, ref3, rtol=1e-13)
    # the lateral ring of the new current arrays
    if wall == "absorbing":
        for a, b in ((gp, pos), (gn, neg)):
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[-1], b[-1])
            np.testing.assert_array_equal(a[:, 0], b[:, 0])
            np.testing.assert_array_equal(a[:, -1], b[:, -1])
    else:
        for a, ref in ((gp, ref_p), (gn, ref_n)):
            m = ref.copy()
            serial.apply_lateral_boundary(m, "reflecting")
            np.testing.assert_array_equal(a[0], m[0])
            np.testing.assert_array_equal(a[-1], m[-1])
            np.testing.assert_array_equal(a[:, 0], m[:, 0])
            np.testing.assert_array_equal(a[:, -1], m[:, -1])


@pytest.mark.parametrize("grid_um,gap_cm", FP32_GRIDS)
@pytest.mark.parametrize("wall", ["absorbing", "reflecting"])
def test_one_fp32_sweep_stays_within_fp32_rounding_of_the_oracle(wall, grid_um, gap_cm):
    config = SimulationConfig(**{**CC, "grid_size_um": grid_um, "electrode_gap_cm": gap_cm}, **TWO_SPECIES,
                              lateral_boundary=wall)
    rng = np.random.default_rng(19)
    shape = (config.no_xy, config.no_xy, config.no_z_with_buffer)
    pos = (rng.random(shape) * 1e7).astype(np.float32).astype(np.float64)
    neg = (rng.random(shape) * 1e7).astype(np.float32).astype(np.float64)
    pos[:, :, [0, -1]] = 0.0
    neg[:, :, [0, -1]] = 0.0
    serial.apply_lateral_boundary(pos, wall)
    serial.apply_lateral_boundary(neg, wall)
    ref_p, ref_n, ref3 = batched.lw_step(config, pos, neg)
    with HipSimulation(config, "fp32") as sim:
        sim.write_fields(pos, neg)
        got3 = sim.step()
        gp, gn = sim.read_fields()
    c = (slice(1, -1), slice(1, -1), slice(1, -1))
    scale = np.abs(ref_p[c]).max()
    assert np.abs(gp[c] - ref_p[c]).max() <= 4e-7 * scale
    assert np.abs(gn[c] - ref_n[c]).max() <= 4e-7 * scale
    np.testing.assert_allclose(got3, ref3, rtol=1e-6)
    # the k end planes stay exactly zero (windows store whole pairs there)
    for a in (gp, gn):
        assert not a[:, :, [0, -1]].any()


@pytest.mark.parametrize("grid_um,gap_cm", FP32_GRIDS)
@pytest.mark.parametrize("wall", ["absorbing", "reflecting"])
def test_fp32_one_two_and_four_values_per_lane_give_the_same_fields(wall, grid_um, gap_cm):
    """The FP32 layouts (4-, 8- and 16-byte lanes in windows of 62, 124 and
    248 values over rows padded to an even length) run the same voxel update:
    a whole run must agree bit for bit, k end planes included. Where the row
    length is 2 mod 4, the four-value lane holding a row's last two values
    loads 8 bytes past the row (the next row, or the slack past the plane and
    the array) and stores those two values only; both walls, so the absorbing
    wall's ring row at the end of each plane is read that way too."""
    from pulsed_ion_chamber_hip.pulses import draw_tracks

    config = SimulationConfig(**{**CC, "grid_size_um": grid_um, "electrode_gap_cm": gap_cm}, **TWO_SPECIES,
                              lateral_boundary=wall)
    draw = draw_tracks(config)
    out = []
    for v in (1, 2, 4):
        with HipSimulation(config, "fp32") as sim:
            sim.set_tiling(values_per_lane=v)
            sim.upload_tracks(*draw)
            sim.run_steps(0, config.total_time_steps)
            out.append((sim.read_per_step(), *sim.read_fields()))
    for other in out[1:]:
        np.testing.assert_array_equal(out[0][1], other[1])
        np.testing.assert_array_equal(out[0][2], other[2])
        # the scored FP64 sums add the same terms in another order (four values
        # per lane fold a lane's values first): rounding of up to ~3e5 terms,
        # n * eps ~ 4e-11 of a sum, and once the gap has emptied the totals are
        # residues of mixed sign far below their peak, so the bound is absolute
        for col in range(out[0][0].shape[1]):
            peak = np.abs(out[0][0][:, col]).max()
            np.testing.assert_allclose(out[0][0][:, col], other[0][:, col], rtol=1e-10, atol=1e-12 * peak)
