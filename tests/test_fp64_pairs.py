"""Two FP64 values per lane.

FP64 grids with enough waves sweep two values per lane (16-byte loads and
stores) in 2-row tiles marching in lockstep pairs. Each voxel's update is the
same arithmetic as with one value per lane, so the fields must equal the
one-value sweep's bit for bit; the scored FP64 sums add the same terms in
another order (a lane's two values are folded first) and agree to rounding.
The grid here has rows of odd length (131 values): FP64 rows are not padded,
so the last lane of a row reads the next row's first value (a k end plane)
and stores a zero there, as that row's own sweep does. Whole runs, so both
the depositing and the clearance kernels run, with both walls; and the
adjoint, whose recomputed segments sweep with the forward tiling."""

import numpy as np
import pytest

from configs import readme_config
from pulsed_ion_chamber_hip import run_simulation_hip
from pulsed_ion_chamber_hip.adjoint import KsModel
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu

ONE_VALUE = (4, 16, 1)


def _odd_rows_config(wall="absorbing"):
    cfg = readme_config(16.0, 0.06, lateral_boundary=wall)
    assert (cfg.no_xy, cfg.no_z_with_buffer) == (83, 131)
    return cfg


def test_default_fp64_tiling_takes_two_values_on_this_grid():
    with HipSimulation(_odd_rows_config(), "fp64") as sim:
        tiling = sim.get_tiling()
    assert (tiling["rows_per_tile"], tiling["values_per_lane"]) == (2, 2)


@pytest.mark.parametrize("wall", ["absorbing", "reflecting"])
def test_two_values_equal_one_value_over_a_whole_run(wall):
    cfg = _odd_rows_config(wall)
    pairs = run_simulation_hip(cfg, progress=False, precision="fp64")
    one = run_simulation_hip(cfg, progress=False, precision="fp64", tiling=ONE_VALUE)
    np.testing.assert_array_equal(pairs.positive_array, one.positive_array)
    np.testing.assert_array_equal(pairs.negative_array, one.negative_array)
    assert pairs.ks == pytest.approx(one.ks, rel=1e-12)
    for name in ("n_positive", "n_negative", "recombination"):
        a, b = np.asarray(getattr(pairs, name)), np.asarray(getattr(one, name))
        np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12 * np.abs(b).max())
    assert np.asarray(one.injected_positive).sum() > 0 and np.any(np.asarray(one.injected_positive) == 0)


@pytest.mark.parametrize("tiling", [(1, 4, 2), (2, 12, 2), (1, 2, 2)], ids=["1-row", "2-row-12", "1-row-2"])
def test_forced_two_value_tilings_equal_one_value(tiling):
    cfg = _odd_rows_config("reflecting")
    a = run_simulation_hip(cfg, progress=False, precision="fp64", tiling=tiling)
    b = run_simulation_hip(cfg, progress=False, precision="fp64", tiling=ONE_VALUE)
    np.testing.assert_array_equal(a.positive_array, b.positive_array)
    np.testing.assert_array_equal(a.negative_array, b.negative_array)


def test_two_values_reject_tall_tiles():
    with HipSimulation(_odd_rows_config(), "fp64") as sim:
        for rows in (3, 4, 8):
            with pytest.raises(ValueError):
                sim.set_tiling(rows, 16, 2)
        sim.set_tiling(4, 0, 0)  # automatic values under 4-row tiles: one
        assert sim.get_tiling()["values_per_lane"] == 1


def test_adjoint_with_two_values_equals_one_value():
    cfg = _odd_rows_config("reflecting")
    ks_pairs, grad_pairs = KsModel(cfg).ks_and_gradient()
    ks_one, grad_one = KsModel(cfg, tiling=ONE_VALUE).ks_and_gradient()
    assert ks_pairs == pytest.approx(ks_one, rel=1e-12)
    np.testing.assert_allclose(grad_pairs, grad_one, rtol=1e-9)
