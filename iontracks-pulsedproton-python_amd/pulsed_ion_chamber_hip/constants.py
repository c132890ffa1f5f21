"""Transport constants of ions in air (Kanai 1998) and related physical values.

The numbers are the same physical inputs the reference package uses
(pulsed_ion_chamber/constants.py:11-36); they are data, so they are kept
identical, and every derived quantity in :mod:`config` depends on them
bit-for-bit.
"""

# eV to create one ion pair in air (proton value), constants.py:11
W_EV_PER_ION_PAIR = 34.2
# single averaged carrier species, constants.py:12-13
ION_MOBILITY_CM2_VS = 1.65
ION_DIFFUSION_CM2_S = 3.7e-2
# volume recombination coefficient, constants.py:14; the solvers multiply it by dt
RECOMBINATION_ALPHA_CM3_S = 1.60e-6

# two resolved Kanai species, constants.py:21-24
ION_MOBILITY_POSITIVE_CM2_VS = 1.36
ION_MOBILITY_NEGATIVE_CM2_VS = 2.10
ION_DIFFUSION_POSITIVE_CM2_S = 2.82e-2
ION_DIFFUSION_NEGATIVE_CM2_S = 4.35e-2

# dry-air densities, constants.py:26-32
AIR_DENSITY_KG_M3 = 1.225
AIR_DENSITY_20C_KG_M3 = 1.2041
JOULE_TO_KEV = 6.241e15
