"""Rows of more k windows than one sweep block holds.

A block holds at most 8 waves (the sweep kernels' 512-thread bound). Rows
whose windows fit stay whole in one block; longer rows are split into groups
of windows, in blocks of two tiles marching in lockstep where every wave of
the block marches and free-running in a block whose group is short. Each
voxel's update is the same arithmetic however the windows and tiles are
grouped, so the fields must equal a reference tiling's bit for bit (the
one-value FP64 sweep; for FP32, one tile per block, or two values per lane
against four); the scored FP64 sums are folded from another partition of
blocks and agree to rounding. Cases: two FP64 values per lane with 5 (groups
2+2+1), 8 (2+2+2+2) and 9 (3+3+3) windows; two FP32 values with 8 windows
(4+4); four FP32 values with 5 windows (2+2+1), which launched blocks of 10
waves before rows were split."""

import numpy as np
import pytest

from configs import readme_config
from pulsed_ion_chamber_hip import _lib
from pulsed_ion_chamber_hip.pulses import draw_tracks
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu

STEPS = 40


def _run(cfg, precision, tiling=None, tiles_per_block=None):
    draw = draw_tracks(cfg)
    with HipSimulation(cfg, precision) as sim:
        if tiling is not None:
            sim.set_tiling(*tiling)
        if tiles_per_block is not None:
            _lib.check(sim.L.pic_set_tile_group(sim.handle, tiles_per_block), "pic_set_tile_group")
        shape = sim.get_tiling()
        sim.upload_tracks(*draw)
        sim.run_steps(0, STEPS)
        return shape, sim.read_per_step(0, STEPS), *sim.read_fields()


# name: (grid, precision, k interior values, (rows, values) of the default
# tiling, reference tiling)
CASES = {
    "fp64-5-windows": ((4.0, 0.02, {}), "fp64", 504, (2, 2), dict(tiling=(4, 16, 1))),
    "fp64-8-windows": ((2.0, 0.01, {"electrode_gap_cm": 0.18}), "fp64", 904, (2, 2), dict(tiling=(4, 16, 1))),
    "fp64-9-windows": ((2.0, 0.01, {}), "fp64", 1004, (2, 2), dict(tiling=(4, 16, 1))),
    "fp32-8-windows": ((2.0, 0.01, {}), "fp32", 1004, (None, 2), dict(tiles_per_block=1)),
    "fp32-four-values-5-windows": ((2.0, 0.012, {"electrode_gap_cm": 0.24}), "fp32", 1204, (3, 4),
                                   dict(tiling=(4, 16, 2))),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_split_rows_equal_the_reference_tiling(case):
    (grid_um, radius, extra), precision, interior, (rows, values), reference = CASES[case]
    cfg = readme_config(grid_um, radius, **extra)
    assert cfg.no_z_with_buffer - 2 == interior
    got = _run(cfg, precision)
    ref = _run(cfg, precision, **reference)
    assert got[0]["values_per_lane"] == values
    if rows is not None:
        assert got[0]["rows_per_tile"] == rows
    np.testing.assert_array_equal(got[2], ref[2])
    np.testing.assert_array_equal(got[3], ref[3])
    rec, rec_ref = got[1], ref[1]
    assert rec_ref[:, 0].sum() > 0  # the steps deposited
    for col in range(rec.shape[1]):
        peak = np.abs(rec_ref[:, col]).max()
        np.testing.assert_allclose(rec[:, col], rec_ref[:, col], rtol=1e-10, atol=1e-12 * peak)
