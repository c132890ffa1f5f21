"""The full Classic Markus electrode at 10 um (536 x 536 x 210, 2194 steps,
24.6 M tracks) on the GPU, FP64 fields, against what is known about it:

* the reference's archived run record (its time and injected-charge columns,
  which this repository's oracle reproduces to 5e-13: the same tracks,
  deposited and scored the same way);
* the oracle's k_s and totals for the same run (computed once, stored with
  the fixture by tests/golden/make_full_electrode_fixture.py, because the
  CPU run takes minutes).
"""

import json
from pathlib import Path

import numpy as np
import pytest

from pulsed_ion_chamber_hip import output, run_simulation_hip
from pulsed_ion_chamber_hip.cli import build_config

pytestmark = pytest.mark.gpu

FIXTURE = json.loads((Path(__file__).parent / "golden" / "full_electrode_10um.json").read_text())


@pytest.fixture(scope="module")
def run():
    cfg = build_config("full_electrode")
    assert [cfg.no_xy, cfg.no_xy, cfg.no_z_with_buffer] == FIXTURE["grid"]
    return run_simulation_hip(cfg, progress=False, return_fields=False)


def test_time_and_injected_columns_match_the_archived_record(run):
    table = output.collected_charge_table(run)
    np.testing.assert_allclose(table["time"], FIXTURE["time"], rtol=1e-12)
    np.testing.assert_allclose(table["injected_positive"], FIXTURE["injected_positive"], rtol=1e-11)
    np.testing.assert_array_equal(table["injected_negative"], table["injected_positive"])


def test_ks_and_totals_match_the_oracle(run):
    ref = FIXTURE["oracle"]
    vol = run.config.unit_length_cm**3
    assert run.ks == pytest.approx(ref["ks"], rel=1e-12)
    assert np.sum(run.injected_positive) * vol == pytest.approx(ref["sum_injected_ion_pairs"], rel=1e-12)
    assert np.sum(run.recombination) * vol == pytest.approx(ref["sum_recombined_ion_pairs"], rel=1e-12)


def test_record_written_by_the_hip_backend_loads_back(run, tmp_path):
    output.save_run_record(run, tmp_path)
    rec = output.load_run_record(tmp_path)
    assert rec.meta["ks"] == run.ks
    assert rec.meta["tracks_per_pulse"] == 24630400
    np.testing.assert_allclose(rec.table["injected_positive"], FIXTURE["injected_positive"], rtol=1e-11)
