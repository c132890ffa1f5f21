"""The run record (the reference's three-file format) and the tier CLI."""

import json

import numpy as np
import pytest

from configs import fast_config
from oracle import serial
from pulsed_ion_chamber_hip import output
from pulsed_ion_chamber_hip.cli import GRID_TIERS, build_config


def test_record_round_trip_in_ion_pairs(tmp_path):
    result = serial.run_serial(fast_config())
    paths = output.save_run_record(result, tmp_path / "run")
    rec = output.load_run_record(tmp_path / "run")
    vol = result.config.unit_length_cm**3
    table = output.collected_charge_table(result)
    for name in output.COLLECTED_CHARGE_COLUMNS:
        np.testing.assert_allclose(rec.table[name], table[name], rtol=1e-12, atol=0)
    np.testing.assert_allclose(rec.table["recombination"], result.recombination * vol, rtol=1e-12)
    np.testing.assert_array_equal(rec.track_density_xy, result.track_density_xy)
    assert rec.meta["ks"] == result.ks and rec.meta["f"] == result.f_t[-1]
    assert rec.config.no_xy == result.config.no_xy
    header = paths["csv"].read_text().splitlines()[0]
    assert header == "time,n_positive,n_negative,injected_positive,injected_negative,recombination"
    first = paths["csv"].read_text().splitlines()[1].split(",")
    assert all("e" in v and len(v.split("e")[0].split(".")[1]) == 12 for v in first)
    # the reference writes with Python's csv module: every line ends in CR LF
    raw = paths["csv"].read_bytes().split(b"\n")
    assert raw[-1] == b"" and all(line.endswith(b"\r") for line in raw[:-1])
    assert len(raw) - 1 == 1 + len(result.time_s)


def test_missing_file_is_named(tmp_path):
    result = serial.run_serial(fast_config())
    output.save_run_record(result, tmp_path / "run")
    (tmp_path / "run" / "run_meta.json").unlink()
    with pytest.raises(FileNotFoundError, match="run_meta.json"):
        output.load_run_record(tmp_path / "run")


def test_cli_tiers_build_the_published_grids():
    # examples/ifj_aic144/README.md tier table: 10 um voxels, nzb 210
    expected = {"dev": 12, "archive": 22, "standard": 34, "wide": 42, "full_electrode": 536}
    for tier, no_xy in expected.items():
        cfg = build_config(tier)
        assert (cfg.no_xy, cfg.no_z_with_buffer) == (no_xy, 210)
    assert build_config("full_electrode").number_of_tracks_per_pulse == 24630400
    assert set(GRID_TIERS) == set(expected)
