"""Two time steps per pass (sweep2_kernel): FP32 fields, either lateral wall.

The pass computes step t+1 on a halo of one row and one k lane each side of
its tile, shares rows through LDS, and then step t+2 on the tile; every voxel
update is the single-step sweep's FP32 arithmetic, so the fields must equal
single steps bit for bit -- odd step counts (the last step alone), runs in
pieces, steps with and without tracks, chunks that do and do not divide the
grid. The per-step records fold the same values in other groupings: the
injected charge is the same gather's, the recombination and totals agree to
rounding."""

import numpy as np
import pytest

from configs import aic_config
from pulsed_ion_chamber_hip.pulses import draw_tracks
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu

# (r cm, grid um, steps): 34^2 x 210, 106^2 x 210, 206^2 x 410 (two k windows)
GRIDS = [(0.014, 10.0, 61), (0.05, 10.0, 40), (0.05, 5.0, 57)]


def run(cfg, draw, steps, pieces=None, steps_per_pass=1, chunk=0, precision="fp32"):
    with HipSimulation(cfg, precision) as sim:
        sim.set_steps_per_pass(steps_per_pass, chunk)
        sim.upload_tracks(*draw)
        first = 0
        for n in pieces or [steps]:
            sim.run_steps(first, n)
            first += n
        return sim.read_per_step(0, steps), sim.read_fields()


def assert_same(single, paired):
    (r1, (p1, n1)), (r2, (p2, n2)) = single, paired
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(n1, n2)
    np.testing.assert_array_equal(r1[:, 0], r2[:, 0])  # the same gather's injected charge
    np.testing.assert_allclose(r1[:, 1], r2[:, 1], rtol=1e-13)  # FP64 sums, other grouping
    # totals: FP64 sums in other groupings; once the gap empties they are
    # residues far below their peak, so the bound is relative to the peak
    for col in (2, 3):
        peak = np.abs(r1[:, col]).max()
        np.testing.assert_allclose(r1[:, col], r2[:, col], rtol=1e-10, atol=1e-12 * peak)


@pytest.mark.parametrize("r_cm, grid_um, steps", GRIDS)
@pytest.mark.parametrize("chunk", [0, 7])
@pytest.mark.parametrize("wall", ["reflecting", "absorbing"])
def test_pairs_equal_single_steps(r_cm, grid_um, steps, chunk, wall):
    """Absorbing: the ring rows and planes carry their values plus both
    steps' depositions; the blocks and chunks at the grid edge store them."""
    cfg = aic_config(r_cm, grid_um, lateral_boundary=wall)
    draw = draw_tracks(cfg)
    single = run(cfg, draw, steps)
    paired = run(cfg, draw, steps, steps_per_pass=2, chunk=chunk)
    assert single[0][:, 0].sum() > 0  # the steps deposited
    assert_same(single, paired)


def test_pairs_in_pieces_equal_one_run():
    """Pieces of odd length leave a step alone at each end; the pending
    records of a pass are folded by the next call's first launch."""
    cfg = aic_config(0.05, 10.0)
    draw = draw_tracks(cfg)
    single = run(cfg, draw, 60)
    assert_same(single, run(cfg, draw, 60, pieces=[7, 1, 2, 13, 37], steps_per_pass=2))


def test_pairs_through_clearance_to_the_end_of_a_run():
    """A whole run (its last steps without tracks): k_s as single steps."""
    from pulsed_ion_chamber_hip import run_simulation_hip
    from pulsed_ion_chamber_hip.state import collection_efficiency

    cfg = aic_config(0.014, 10.0)
    draw = draw_tracks(cfg)
    total = cfg.total_time_steps
    single = run(cfg, draw, total)
    paired = run(cfg, draw, total, steps_per_pass=2)
    assert_same(single, paired)
    ks = [1.0 / collection_efficiency(r[:, 0], r[:, 1], total)[-1] for r, _ in (single, paired)]
    assert ks[1] == pytest.approx(ks[0], rel=1e-13)
    assert run_simulation_hip(cfg, progress=False, precision="fp32").ks == pytest.approx(ks[0], rel=1e-13)


def test_the_library_pairs_large_fp32_grids_only():
    """By default, 206^2 x 410 (FP32, either wall) takes pairs: half as many
    sweep launches as steps; forcing a single-step tiling, or FP64 fields,
    gives one launch per step."""
    cfg = aic_config(0.05, 5.0)
    draw = draw_tracks(cfg)

    def launches(precision="fp32", tiling=None, config=cfg):
        with HipSimulation(config, precision) as sim:
            if tiling:
                sim.set_tiling(*tiling)
            sim.upload_tracks(*draw)
            sim.enable_timing(True)
            sim.run_steps(0, 40)
            return sim.timing()["sweep_launches"]

    assert launches() == 20
    assert launches(tiling=(3, 16, 4)) == 40
    assert launches(precision="fp64") == 40
    assert launches(config=aic_config(0.05, 5.0, lateral_boundary="absorbing")) == 20


def test_pairs_are_refused_where_the_pass_does_not_exist():
    with HipSimulation(aic_config(0.05, 10.0), "fp64") as sim:
        with pytest.raises(ValueError, match="FP32"):
            sim.set_steps_per_pass(2)
    with HipSimulation(aic_config(0.05, 10.0, lateral_boundary="absorbing"), "fp32") as sim:
        sim.set_steps_per_pass(2)
        with pytest.raises(ValueError):
            sim.set_steps_per_pass(3)


def test_pairs_replayed_through_graph_chains_equal_direct_launches():
    """With graph replay, a segment's pass launches (two steps each, their
    folds in the next launch's first blocks) are collected and replayed
    through the chains: the same fields and records."""
    cfg = aic_config(0.05, 5.0)
    draw = draw_tracks(cfg)
    out = []
    for graphs in (False, True):
        with HipSimulation(cfg, "fp32") as sim:
            sim.set_graphs(graphs)
            sim.upload_tracks(*draw)
            sim.run_steps(0, 45)
            sim.run_steps(45, 30)
            out.append((sim.read_per_step(0, 75), sim.read_fields()))
    np.testing.assert_array_equal(out[0][1][0], out[1][1][0])
    np.testing.assert_array_equal(out[0][1][1], out[1][1][1])
    np.testing.assert_array_equal(out[0][0], out[1][0])
