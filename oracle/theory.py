"""Jaffe's closed-form initial-recombination k_s for one track. Test
infrastructure only (the single-track acceptance test compares the simulated
k_s with it).

Restates pulsed_ion_chamber/theory.py:43-76 with mpmath's exponential
integral at 50 significant digits.
"""

from math import log, pi

import mpmath

from pulsed_ion_chamber_hip.constants import ION_DIFFUSION_CM2_S, ION_MOBILITY_CM2_VS, W_EV_PER_ION_PAIR
from pulsed_ion_chamber_hip.stopping_power import calc_track_radius_cm

ALPHA_CM3_S = 1.60e-6


def jaffe_ks(LET_keV_um, voltage_V, electrode_gap_cm, W_eV=None, mu=None, D=None, alpha=None):
    W_eV = W_EV_PER_ION_PAIR if W_eV is None else W_eV
    mu = ION_MOBILITY_CM2_VS if mu is None else mu
    D = ION_DIFFUSION_CM2_S if D is None else D
    alpha = ALPHA_CM3_S if alpha is None else alpha
    field = voltage_V / electrode_gap_cm
    b = calc_track_radius_cm(LET_keV_um)
    n0 = LET_keV_um * 1e7 / W_eV
    g = alpha * n0 / (8.0 * pi * D)
    with mpmath.workdps(50):
        factor = mpmath.exp(-1.0 / g) * mu * b**2 * field / (2.0 * g * electrode_gap_cm * D)
        upper = mpmath.ei(1.0 / g + log(1.0 + (2.0 * electrode_gap_cm * D / (mu * b**2 * field))))
        lower = mpmath.ei(1.0 / g)
        f = factor * (upper - lower)
        return float(1.0 / f)
