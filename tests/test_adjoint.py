"""The GPU reverse pass: dk_s/d(alpha, mu+, mu-, D+, D-, track radius).

Checked three ways: against the CPU reverse-mode oracle (itself pinned to
finite differences in test_adjoint_oracle.py), against FP64 central finite
differences of the GPU forward run (relative step 1e-6, dt and step count
frozen), and for independence from the checkpoint interval and from the
number of checkpoint levels (the recomputed segments reproduce the forward
states exactly, so every choice must give the same numbers bit for bit).
"""

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES, fast_config
from oracle import adjoint as oracle_adjoint
from pulsed_ion_chamber_hip import SimulationConfig
from pulsed_ion_chamber_hip.adjoint import PARAMS, KsModel, ks_torch
from pulsed_ion_chamber_hip.solver_hip import HipSimulation

pytestmark = pytest.mark.gpu

CASES = {
    "absorbing": dict(SMALL),
    "reflecting-two-species": {**SMALL, **TWO_SPECIES, "lateral_boundary": "reflecting"},
    "absorbing-two-species-full-grid": {**SMALL, **TWO_SPECIES, "scoring_region": "full_grid", "seed": 5},
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_gradient_matches_the_oracle(case):
    cfg = SimulationConfig(**CASES[case])
    ks, grad = KsModel(cfg).ks_and_gradient()
    fz = oracle_adjoint.freeze(cfg)
    ks_ref, grad_ref, _ = oracle_adjoint.gradient(fz, oracle_adjoint.base_theta(cfg))
    assert ks == pytest.approx(ks_ref, rel=1e-12)
    np.testing.assert_allclose(grad, grad_ref, rtol=1e-10)


def test_gpu_gradient_matches_gpu_finite_differences():
    cfg = SimulationConfig(**{**SMALL, **TWO_SPECIES, "lateral_boundary": "reflecting"})
    model = KsModel(cfg)
    _, grad = model.ks_and_gradient()
    fd = model.finite_difference()
    np.testing.assert_allclose(grad, fd, rtol=2e-7)


def test_gradient_on_three_pulses_matches_richardson_finite_differences():
    """The adjoint benchmark configuration: 10 um voxels over r 0.02 cm
    (48^2 x 206), three pulses (106 166 steps), FP64, checkpoints every 64
    steps. All six derivatives agree with Richardson-extrapolated central
    differences of the GPU forward run (relative steps 1e-3 and 2e-3) to
    1e-6 relative; a plain quotient at step 1e-6 resolves only ~1e-4 here
    (the rounding noise of k_s over 1e5 steps, divided by the step)."""
    from configs import readme_config

    cfg = readme_config(10.0, 0.02, n_pulses=3)
    assert (cfg.no_xy, cfg.no_z_with_buffer, cfg.total_time_steps) == (48, 206, 106166)
    model = KsModel(cfg, checkpoint_every=64)
    ks, grad = model.ks_and_gradient()
    fd = model.finite_difference_richardson(rel=1e-3)
    rel = np.abs(grad - fd) / np.abs(fd)
    print("3-pulse adjoint vs Richardson FD, relative error per parameter:", dict(zip(PARAMS, rel.tolist())))
    assert ks == pytest.approx(model.ks(model.theta0), rel=1e-14)
    assert np.all(rel <= 1e-6), rel


def test_forward_at_theta0_is_the_plain_run():
    from pulsed_ion_chamber_hip import run_simulation_hip

    cfg = fast_config()
    model = KsModel(cfg)
    assert model.ks(model.theta0) == run_simulation_hip(cfg, progress=False).ks


@pytest.mark.parametrize("every", [1, 5, 64])
def test_checkpoint_interval_does_not_change_the_result(every):
    cfg = SimulationConfig(**{**SMALL, **TWO_SPECIES})
    base = KsModel(cfg, checkpoint_every=16).ks_and_gradient()
    got = KsModel(cfg, checkpoint_every=every).ks_and_gradient()
    assert got[0] == base[0]
    np.testing.assert_array_equal(got[1], base[1])


@pytest.mark.parametrize("outer", [0, 1, 3, 7, 1000])
def test_two_level_checkpoints_do_not_change_the_result(outer):
    """Checkpoints every step give hundreds of segments; keeping only every
    M-th segment's start and recomputing the spans in the reverse pass (M = 3
    and 7 leave a short last span; 1000 is one span over the whole run; 0
    picks M = ceil(sqrt(segments))) must give the one-level numbers bit for
    bit, and hold fewer states."""
    cfg = SimulationConfig(**{**SMALL, **TWO_SPECIES})
    one = KsModel(cfg, checkpoint_every=1, segments_per_outer=1)
    base = one.ks_and_gradient()
    assert one.last_info["segments_per_outer"] == 1
    nseg = one.last_info["segments"]
    assert nseg == cfg.total_time_steps > 64
    model = KsModel(cfg, checkpoint_every=1, segments_per_outer=outer)
    got = model.ks_and_gradient()
    info = model.last_info
    m = {0: int(np.ceil(np.sqrt(nseg))), 1000: nseg}.get(outer, outer)
    assert info["segments_per_outer"] == m
    assert info["outer_checkpoints"] == -(-nseg // m)
    if 1 < m < nseg:
        assert info["device_bytes"] < one.last_info["device_bytes"]
    assert got[0] == base[0]
    np.testing.assert_array_equal(got[1], base[1])


def test_torch_function_backward_is_the_adjoint():
    import torch

    cfg = SimulationConfig(**SMALL)
    model = KsModel(cfg)
    theta = torch.tensor(model.theta0, dtype=torch.float64, requires_grad=True)
    ks = ks_torch(model, theta)
    (2.0 * ks).backward()
    _, grad = model.ks_and_gradient()
    np.testing.assert_allclose(theta.grad.numpy(), 2.0 * grad, rtol=1e-14)
    with torch.no_grad():
        assert float(ks_torch(model, theta.detach())) == pytest.approx(float(ks), rel=1e-14)


def test_adjoint_rejects_fp32_handles():
    import ctypes

    from pulsed_ion_chamber_hip import _lib

    cfg = SimulationConfig(**SMALL)
    with HipSimulation(cfg, "fp32") as sim:
        sim.upload_tracks(np.array([1] + [0] * (cfg.total_time_steps - 1)), np.array([2.0]), np.array([2.0]))
        adj = ctypes.c_void_p()
        with pytest.raises(ValueError, match="FP64"):
            _lib.check(_lib.open_library().pic_adjoint_create(sim.handle, 16, ctypes.byref(adj)), "pic_adjoint_create")
