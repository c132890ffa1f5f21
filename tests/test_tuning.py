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
    assert tuple(first["tiling"]) in tuning.candidates(precision)
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
