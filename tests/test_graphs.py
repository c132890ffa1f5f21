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
    """A batch of four configurations, with and without graph replay: the
    same two lanes (the longest run on its own host thread) launch the same
    kernels, so fields and per-step records agree bit for bit."""
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
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_two_threads_replay_graphs_on_their_own_handles():
    """Two handles with graph replay driven from two host threads at once
    (each ctypes call releases the GIL, so both threads are inside the
    library together and share the device's pool of graph chains): every
    run, repeated, equals the same run made with direct launches."""
    import threading

    cfgs = [readme_config(16.0, 0.0077, voltage_V=69.47, dose_rate_Gy_s=40.0),
            readme_config(12.5, 0.0052, voltage_V=134.13, dose_rate_Gy_s=40.0)]
    want = [_run_single(c, "fp32", graphs=False) for c in cfgs]
    got = [[], []]
    errors = []

    def drive(k):
        try:
            for _ in range(4):
                got[k].append(_run_single(cfgs[k], "fp32", graphs=True))
        except Exception as exc:  # reported by the main thread
            errors.append(exc)

    threads = [threading.Thread(target=drive, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(2):
        assert len(got[k]) == 4
        for run in got[k]:
            for a, b in zip(want[k], run):
                np.testing.assert_array_equal(a, b)
