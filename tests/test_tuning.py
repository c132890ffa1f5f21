"""Autotuned sweep tiling: measured, cached, and without effect on results."""

import json

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES
from pulsed_ion_chamber_hip import SimulationConfig, run_simulation_hip
from pulsed_ion_chamber_hip import tuning

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_measures_caches_and_keeps_the_fields(tmp_path, monkeypatch, precision):
    monkeypatch.setenv("PIC_HIP_TUNING_CACHE", str(tmp_path / "tiling.json"))
    cfg = SimulationConfig(**{**SMALL, **TWO_SPECIES, "grid_size_um": 20.0})
    first = tuning.autotune(cfg, precision, steps=12, repeats=1)
    assert not first["cached"]
    # the winner is a candidate or the library's own choice, which competes
    assert tuple(first["tiling"]) in tuning.candidates(precision) + [tuple(first["default_tiling"])]
    assert first["measured"][",".join(map(str, first["default_tiling"]))] == first["default_us"]
    assert first["sweep_us"] == min(first["measured"].values())
    table = json.loads((tmp_path / "tiling.json").read_text())
    assert tuning.cache_key(cfg, precision) in table
    again = tuning.autotune(cfg, precision, steps=12, repeats=1)
    assert again["cached"] and again["tiling"] == first["tiling"]
    assert tuning.cached_tiling(cfg, precision) == tuple(first["tiling"])
    tuned = run_simulation_hip(cfg, progress=False, precision=precision, tiling="auto")
    plain = run_simulation_hip(cfg, progress=False, precision=precision)
    np.testing.assert_array_equal(tuned.positive_array, plain.positive_array)
    np.testing.assert_array_equal(tuned.negative_array, plain.negative_array)
    assert tuned.ks == pytest.approx(plain.ks, rel=1e-12)


def test_explicit_tiling_and_bad_values():
    cfg = SimulationConfig(**SMALL)
    a = run_simulation_hip(cfg, progress=False, tiling=(1, 4, 1))
    b = run_simulation_hip(cfg, progress=False, tiling=(2, 16, 1))
    np.testing.assert_array_equal(a.positive_array, b.positive_array)
    with pytest.raises(ValueError):
        run_simulation_hip(cfg, progress=False, tiling="fast")
    with pytest.raises(ValueError):
        run_simulation_hip(cfg, progress=False, tiling=(3, 4, 1))


@pytest.mark.parametrize("precision, values", [("fp32", 2), ("fp32", 1), ("fp64", 0)])
@pytest.mark.parametrize("wall", ["absorbing", "reflecting"])
def test_short_march_matches_the_general_march(precision, values, wall):
    """1-row tiles of one or two planes run the short march (exactly that many
    plane steps); four planes the general one. Same fields bit for bit."""
    from configs import readme_config

    cfg = readme_config(20.0, 0.0101, voltage_V=96.53, dose_rate_Gy_s=40.0, lateral_boundary=wall)
    runs = [run_simulation_hip(cfg, progress=False, precision=precision, tiling=(1, chunk, values)) for chunk in (4, 2, 1)]
    for other in runs[1:]:
        np.testing.assert_array_equal(other.positive_array, runs[0].positive_array)
        np.testing.assert_array_equal(other.negative_array, runs[0].negative_array)
        assert other.ks == pytest.approx(runs[0].ks, rel=1e-12)
    assert runs[0].injected_positive.sum() > 0


def test_default_tiling_follows_the_grid():
    """Four FP32 values per lane (3-row tiles) and two FP64 values (2-row
    tiles) only on grids with enough waves for them; automatic values fall
    back to two FP32 values and one FP64 value under forced 4-row tiles; the
    default four-value run matches two values per lane bit for bit."""
    from configs import readme_config
    from pulsed_ion_chamber_hip.pulses import draw_tracks
    from pulsed_ion_chamber_hip.solver_hip import HipSimulation

    small = readme_config(16.0, 0.0077)
    large = readme_config(5.0, 0.065)  # 268^2 x 406
    with HipSimulation(small, "fp32") as sim:
        assert sim.get_tiling()["values_per_lane"] == 2
    with HipSimulation(small, "fp64") as sim:
        assert sim.get_tiling()["values_per_lane"] == 1
    with HipSimulation(large, "fp64") as sim:
        tiling = sim.get_tiling()
        assert (tiling["rows_per_tile"], tiling["values_per_lane"]) == (2, 2)
        sim.set_tiling(4, 16, 0)
        assert sim.get_tiling() == dict(rows_per_tile=4, planes_per_chunk=16, values_per_lane=1)
    draw = draw_tracks(large)
    out = []
    with HipSimulation(large, "fp32") as sim:
        tiling = sim.get_tiling()
        assert (tiling["rows_per_tile"], tiling["values_per_lane"]) == (3, 4)
        sim.set_tiling(4, 16, 0)
        assert sim.get_tiling() == dict(rows_per_tile=4, planes_per_chunk=16, values_per_lane=2)
        sim.set_tiling(0, 0, 0)
        assert sim.get_tiling() == tiling
        sim.upload_tracks(*draw)
        for t in ((0, 0, 0), (4, 16, 2)):
            sim.set_tiling(*t)
            sim.reset()
            sim.run_steps(0, 30)
            out.append((sim.read_per_step(0, 30), *sim.read_fields()))
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])
    # the scored FP64 sums add the same terms in another order (a lane's four
    # values are folded first): rounding only
    for col in range(out[0][0].shape[1]):
        peak = np.abs(out[0][0][:, col]).max()
        np.testing.assert_allclose(out[0][0][:, col], out[1][0][:, col], rtol=1e-10, atol=1e-12 * peak)
    assert out[0][0][:, 0].sum() > 0  # the steps deposited


def test_cached_tiling_the_library_rejects_is_measured_again(tmp_path, monkeypatch):
    """A cache entry this library does not accept (written by another build
    under the same key) is replaced by a fresh measurement, not raised."""
    monkeypatch.setenv("PIC_HIP_TUNING_CACHE", str(tmp_path / "tiling.json"))
    cfg = SimulationConfig(**{**SMALL, "grid_size_um": 20.0})
    key = tuning.cache_key(cfg, "fp64")
    (tmp_path / "tiling.json").write_text(json.dumps({key: {"tiling": [3, 4, 1], "sweep_us": 1.0}}))
    res = run_simulation_hip(cfg, progress=False, tiling="auto")
    plain = run_simulation_hip(cfg, progress=False)
    np.testing.assert_array_equal(res.positive_array, plain.positive_array)
    entry = json.loads((tmp_path / "tiling.json").read_text())[key]
    assert tuple(entry["tiling"]) in tuning.candidates("fp64") + [tuple(entry["default_tiling"])]


def test_cache_key_names_the_library_build():
    from pulsed_ion_chamber_hip import _lib

    cfg = SimulationConfig(**SMALL)
    assert f"lib {_lib.library_digest()}" in tuning.cache_key(cfg, "fp64")


def test_one_round_grids_take_the_shortest_chunk_that_fits_one_round():
    """The mid-size 208^2 x 406 grid: four-value 3-row pairs whose blocks all
    fit one round of resident waves at 15 planes (490 blocks of 4 waves for
    the 2048 waves the 256 CUs of an MI355X hold) but not at 14 (525): the
    library takes 15, between the powers of two, and the fields equal those
    of 16-plane chunks bit for bit."""
    from configs import readme_config
    from pulsed_ion_chamber_hip.pulses import draw_tracks
    from pulsed_ion_chamber_hip.solver_hip import HipSimulation

    cfg = readme_config(5.0, 0.05)
    draw = draw_tracks(cfg)
    out = []
    with HipSimulation(cfg, "fp32") as sim:
        assert sim.get_tiling() == dict(rows_per_tile=3, planes_per_chunk=15, values_per_lane=4)
        sim.upload_tracks(*draw)
        for t in ((0, 0, 0), (3, 16, 4)):
            sim.set_tiling(*t)
            sim.reset()
            sim.run_steps(0, 40)
            out.append(sim.read_fields())
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
