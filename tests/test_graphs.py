"""Graph replay of a run's launches must not change a single bit.

With graph replay on, a run's launches (each gather batch's deposition gather and its
per-step sweeps) are collected and replayed through graphs of kernel nodes
that are re-parameterised for every use -- kernel, grid and argument -- so
one node carries a depositing sweep in one use and a plain sweep (another
kernel) in the next. The same launches made one at a time are the reference:
fields, per-step records and the batched path must agree exactly. Runs are
long enough (hundreds of steps) that every graph instance in the pool is
re-parameterised several times, and the first run of a fresh handle follows
a run of another grid so nodes change grid size as well.
"""

import numpy as np
import pytest

from configs import readme_config
from pulsed_ion_chamber_hip.batch import HipBatch
from pulsed_ion_chamber_hip.pulses import draw_tracks
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu


def _run_single(cfg, precision, graphs, first_count=None):
    draw = draw_tracks(cfg)
    with HipSimulation(cfg, precision) as sim:
        sim.set_graphs(graphs)
        sim.upload_tracks(*draw)
        total = cfg.total_time_steps
        # two calls: a run split at an arbitrary step must equal one call
        split = first_count if first_count is not None else total // 3
        sim.run_steps(0, split)
        sim.run_steps(split, total - split)
        sim.synchronize()
        rec = sim.read_per_step()
        pos, neg = sim.read_fields()
    return rec, pos, neg


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("wall", ["absorbing", "reflecting"])
def test_graph_replay_matches_direct_launches(precision, wall):
    cfg = readme_config(16.0, 0.0077, voltage_V=69.47, dose_rate_Gy_s=40.0, lateral_boundary=wall)
    assert cfg.total_time_steps > 200  # several 32-node chains and a remainder
    direct = _run_single(cfg, precision, graphs=False)
    replay = _run_single(cfg, precision, graphs=True)
    for a, b in zip(direct, replay):
        np.testing.assert_array_equal(a, b)
    # the run deposits on some steps and not on others: both kernels were replayed
    assert np.any(direct[0][:, 0] > 0) and np.any(direct[0][:, 0] == 0)


def test_graph_replay_after_another_grid():
    """The pool's graphs were last used by a different grid (other kernels,
    other grid sizes); replaying this one through them must still match."""
    other = readme_config(20.0, 0.0146, voltage_V=50.0, dose_rate_Gy_s=2.0)
    _run_single(other, "fp32", graphs=True)
    cfg = readme_config(12.5, 0.0052, voltage_V=134.13, dose_rate_Gy_s=40.0)
    direct = _run_single(cfg, "fp32", graphs=False, first_count=77)
    replay = _run_single(cfg, "fp32", graphs=True, first_count=77)
    for a, b in zip(direct, replay):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_graph_replay_matches_direct_launches(precision):
    """A batch with graph replay runs as one lane (graph chains are driven
    from one host thread), the direct run of four configurations as two: the
    fields are the same bit for bit, the per-step sums are folded over another
    partition of blocks and agree to rounding."""
    cfgs = [
        readme_config(20.0, 0.0146, voltage_V=50.0, dose_rate_Gy_s=2.0),
        readme_config(16.0, 0.0077, voltage_V=69.47, dose_rate_Gy_s=2.0),
        readme_config(20.0, 0.0101, voltage_V=96.53, dose_rate_Gy_s=40.0),
        readme_config(12.5, 0.0052, voltage_V=134.13, dose_rate_Gy_s=40.0),
    ]
    draws = [draw_tracks(c) for c in cfgs]
    out = {}
    for graphs in (False, True):
        sims = [HipSimulation(c, precision) for c in cfgs]
        try:
            for s, d in zip(sims, draws):
                s.upload_tracks(*d)
            with HipBatch(sims) as b:
                b.set_graphs(graphs)
                b.run()
                b.synchronize()
            out[graphs] = [(s.read_per_step(), *s.read_fields()) for s in sims]
        finally:
            for s in sims:
                s.close()
    for a, b in zip(out[False], out[True]):
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(a[2], b[2])
        for col in range(a[0].shape[1]):
            peak = np.abs(a[0][:, col]).max()
            np.testing.assert_allclose(b[0][:, col], a[0][:, col], rtol=1e-13, atol=1e-14 * peak)
