"""The tier CLI on the GPU: sizing, a short estimate, and a full small run
that writes the record and JSON summary."""

import json

import pytest

from pulsed_ion_chamber_hip import cli, output

pytestmark = pytest.mark.gpu


def test_dry_run_prints_the_sizing(capsys):
    assert cli.main(["full_electrode", "--dry-run"]) == 0
    out = capsys.readouterr().out
    assert "536 x 536 x 210" in out and "No simulation was run" in out


def test_estimate_runtime(capsys):
    assert cli.main(["dev", "--estimate-runtime-seconds", "0.5"]) == 0
    assert "estimated total loop time" in capsys.readouterr().out


def test_dev_tier_run_writes_record_and_json(tmp_path, capsys):
    js = tmp_path / "run.json"
    assert cli.main(["dev", "--save-run", str(tmp_path / "rec"), "--json", str(js)]) == 0
    summary = json.loads(js.read_text())
    rec = output.load_run_record(tmp_path / "rec")
    assert summary["ks"] == pytest.approx(rec.meta["ks"])
    assert (summary["no_xy"], summary["no_z_with_buffer"]) == (12, 210)
    # the reference's documented dev-tier value (tests/test_examples_cli.py:53)
    assert summary["ks"] == pytest.approx(1.0580, abs=1e-4)
