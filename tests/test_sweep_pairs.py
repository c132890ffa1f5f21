"""Sweep blocks that hold two tiles.

Large grids of FP32 fields run two 3-row tiles per block, marching their
planes in lockstep (one barrier per plane, taken only in blocks whose waves
all march). Each voxel's update is the same arithmetic however the tiles are
grouped, so fields must equal the one-tile-per-block sweep bit for bit; the
scored sums are folded from another partition of blocks and agree to
rounding. Both walls, depositing and clearance steps, an even number of tiles
(every block a full pair) and an odd one (the last block holds one tile and
marches without barriers)."""

import ctypes

import numpy as np
import pytest

from configs import aic_config
from pulsed_ion_chamber_hip import _lib
from pulsed_ion_chamber_hip.pulses import draw_tracks
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu


def _run(cfg, tiles_per_block, steps):
    draw = draw_tracks(cfg)
    with HipSimulation(cfg, "fp32") as sim:
        _lib.check(sim.L.pic_set_tile_group(sim.handle, tiles_per_block), "pic_set_tile_group")
        tiling = sim.get_tiling()
        sim.upload_tracks(*draw)
        sim.run_steps(0, steps)
        return tiling, sim.read_per_step(0, steps), *sim.read_fields()


@pytest.mark.parametrize("wall", ["reflecting", "absorbing"])
@pytest.mark.parametrize("radius", [0.0995, 0.1], ids=["even-tiles", "odd-tiles"])
def test_two_tiles_per_block_equal_one(wall, radius):
    cfg = aic_config(radius, 5.0, lateral_boundary=wall)
    rows = cfg.no_xy - 2
    # 402 rows: 134 full 3-row tiles; 404 rows: 135 tiles, the last of two rows
    assert -(-rows // 3) % 2 == (0 if radius == 0.0995 else 1)
    steps = cfg.total_time_steps  # the pulse (3706 depositing steps) and the clearance after it
    one = _run(cfg, 1, steps)
    two = _run(cfg, 0, steps)  # automatic: two tiles per block
    assert one[0] == two[0] and two[0]["values_per_lane"] == 4 and two[0]["rows_per_tile"] == 3
    np.testing.assert_array_equal(one[2], two[2])
    np.testing.assert_array_equal(one[3], two[3])
    rec1, rec2 = one[1], two[1]
    assert rec1[:, 0].sum() > 0 and np.any(rec1[:, 0] == 0)
    for col in range(4):
        peak = np.abs(rec1[:, col]).max()
        np.testing.assert_allclose(rec2[:, col], rec1[:, col], rtol=1e-12, atol=1e-13 * peak)
