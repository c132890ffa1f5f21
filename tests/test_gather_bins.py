"""Deposition with the tracks binned by tile against deposition by scanning.

Each gather tile needs the tracks whose stencil box meets it, in ascending
order (the order of the reference's CSR row accumulation,
solver_numba_parallel.py:140-219). The library finds them either by testing
every track of the step against the tile, or by binning the tracks by tile on
the device first (count, scan, fill; each tile then sorts its bin). Both must
give the same density, the same scored injected charge and so the same run,
bit for bit: fields, per-step records, on small and large grids, both walls,
with and without a track cutoff, and where a bin outgrows what a block sorts
in LDS (2048 tracks) and the block falls back to the scan; bins of up to 256
tracks are sorted by rank, larger ones by a bitonic network.
"""

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES, aic_config, readme_config
from pulsed_ion_chamber_hip import SimulationConfig
from pulsed_ion_chamber_hip.pulses import draw_tracks
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu

CASES = {
    "small-absorbing": (SimulationConfig(**SMALL), None),
    "small-reflecting-two-species": (SimulationConfig(**SMALL, **TWO_SPECIES, lateral_boundary="reflecting"), None),
    "small-no-cutoff": (SimulationConfig(**SMALL, track_cutoff_sigmas=None), None),
    "mid-size-5um": (readme_config(5.0, 0.05), 300),
    "full-electrode-10um": (aic_config(0.265, 10.0), 160),
    # ~1200 tracks per in-pulse step on 2 x 2 tiles: bins of 256..2048 (bitonic sort)
    "large-bins": (readme_config(20.0, 0.0101, dose_rate_Gy_s=200.0), 120),
    # ~6000 tracks per in-pulse step on 2 x 2 tiles: every bin passes 2048
    "overfull-bins": (readme_config(20.0, 0.0101, dose_rate_Gy_s=1000.0), 120),
}


def _run(cfg, draw, steps, mode, precision):
    with HipSimulation(cfg, precision) as sim:
        sim.set_gather_mode(mode)
        sim.upload_tracks(*draw)
        assert sim.gather_binned() == (mode == "binned")
        sim.run_steps(0, steps)
        return sim.read_per_step(0, steps), *sim.read_fields()


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_binned_deposition_equals_the_scan(case, precision):
    cfg, steps = CASES[case]
    steps = cfg.total_time_steps if steps is None else steps
    draw = draw_tracks(cfg)
    if case == "overfull-bins":
        assert draw[0][:steps].min() > 2048
    if case == "large-bins":
        assert 256 < draw[0][:steps].min() and draw[0][:steps].max() <= 2048
    scan = _run(cfg, draw, steps, "scan", precision)
    binned = _run(cfg, draw, steps, "binned", precision)
    assert scan[0][:, 0].sum() > 0
    for a, b in zip(scan, binned):
        np.testing.assert_array_equal(a, b)


def test_automatic_mode_bins_large_grids_only():
    small = readme_config(5.0, 0.05)       # 208^2: 169 tiles
    large = aic_config(0.265, 10.0)        # 536^2: 1156 tiles
    for cfg, want in ((small, False), (large, True)):
        with HipSimulation(cfg, "fp32") as sim:
            sim.upload_tracks(*draw_tracks(cfg))
            assert sim.gather_binned() is want
            sim.set_gather_mode("scan")
            assert not sim.gather_binned()
            sim.set_gather_mode("binned")
            assert sim.gather_binned()


def test_mode_set_before_upload_applies_to_it():
    cfg = SimulationConfig(**SMALL)
    with HipSimulation(cfg, "fp64") as sim:
        sim.set_gather_mode("binned")
        assert not sim.gather_binned()  # nothing uploaded yet
        sim.upload_tracks(*draw_tracks(cfg))
        assert sim.gather_binned()
        with pytest.raises(ValueError):
            sim.set_gather_mode("fastest")
