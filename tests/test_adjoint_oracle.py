"""The CPU reverse-mode oracle for dk_s/dtheta, pinned by mathematics.

Its forward model must be the reference's loop (k_s equal to the C/OpenMP
restatement of the batched backend), and its reverse pass must agree with
FP64 central finite differences of that same model (relative step 1e-6) on
all six parameters, under both walls and with one or two carrier species.
A wrong transposition (z weights not exchanged, the ring not carried or not
folded, a dropped recombination or deposition term) moves at least one
component by far more than the tolerance.
"""

import numpy as np
import pytest

from configs import SMALL, TWO_SPECIES, fast_config
from oracle import adjoint, batched
from pulsed_ion_chamber_hip import SimulationConfig

CASES = {
    "absorbing": dict(SMALL),
    "reflecting-two-species": {**SMALL, **TWO_SPECIES, "lateral_boundary": "reflecting"},
    "absorbing-two-species-full-grid": {**SMALL, **TWO_SPECIES, "scoring_region": "full_grid", "seed": 5},
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_model_forward_is_the_reference_loop(case):
    cfg = SimulationConfig(**CASES[case])
    fz = adjoint.freeze(cfg)
    _, _, ks, _ = adjoint.forward(fz, adjoint.base_theta(cfg))
    assert ks == pytest.approx(batched.run_batched(cfg, fields=False).ks, rel=1e-13)


@pytest.mark.parametrize("case", sorted(CASES))
def test_reverse_pass_matches_central_differences(case):
    cfg = SimulationConfig(**CASES[case])
    fz = adjoint.freeze(cfg)
    theta = adjoint.base_theta(cfg)
    ks, grad, _ = adjoint.gradient(fz, theta)
    fd = adjoint.finite_difference(fz, theta)
    np.testing.assert_allclose(grad, fd, rtol=2e-7)
    assert np.all(np.abs(grad * theta) > 1e-4)  # every parameter matters here


def test_small_gradients_agree_to_the_difference_quotients_noise():
    # k_s - 1 ~ 1e-3 on this grid: finite differences resolve ~1e-9 of k_s
    cfg = fast_config()
    fz = adjoint.freeze(cfg)
    theta = adjoint.base_theta(cfg)
    _, grad, _ = adjoint.gradient(fz, theta)
    fd = adjoint.finite_difference(fz, theta)
    np.testing.assert_allclose(grad * theta, fd * theta, rtol=0, atol=5e-10)


def test_a_wrong_transpose_is_caught():
    """Plant the classic mistake -- transposing the stencil without
    exchanging its z weights -- and watch the mobility derivatives fail."""
    cfg = SimulationConfig(**SMALL)
    fz = adjoint.freeze(cfg)
    theta = adjoint.base_theta(cfg)
    fd = adjoint.finite_difference(fz, theta)
    real = adjoint.weights

    def swapped(f, diffusion, mobility):
        (lat, zm, zp, cen), sc = real(f, diffusion, mobility)
        return (lat, zp, zm, cen), sc

    # the forward pass must stay right, so swap only while the reverse pass runs
    fwd = adjoint.forward(fz, theta, store=True)
    adjoint.weights = swapped
    try:
        orig_forward = adjoint.forward
        adjoint.forward = lambda *a, **k: fwd
        _, bad, _ = adjoint.gradient(fz, theta)
    finally:
        adjoint.weights = real
        adjoint.forward = orig_forward
    assert not np.allclose(bad[1:3], fd[1:3], rtol=1e-3)
