"""LET lookup, Gaussian track radius and dose-rate to fluence-rate conversion.

Follows pulsed_ion_chamber/stopping_power.py:18-56. The same library calls are
used on purpose (pandas CSV parsing, scipy linear interpolation, numpy
quadratic fit): the track count is ``int(round(...))`` of a product that
contains the LET, so a last-bit difference in parsing or interpolation could
move an integer that the whole run depends on.
"""

from functools import lru_cache
from pathlib import Path

import numpy as np

from .constants import AIR_DENSITY_KG_M3, JOULE_TO_KEV

DATA_DIR = Path(__file__).parent / "data"


@lru_cache(maxsize=None)
def _let_table(particle: str):
    import pandas as pd

    table = pd.read_csv(DATA_DIR / "stopping_power_air.csv", skiprows=3)
    column = f"{particle}_LET_keV_um"
    if column not in table.columns:
        raise ValueError(f"Particle '{particle}' not found in stopping power table")
    return table["E_MeV_u"].to_numpy(), table[column].to_numpy()


def E_MeV_u_to_LET_keV_um(E_MeV_u, particle="proton"):
    """Linear interpolation of the tabulated dry-air LET (stopping_power.py:18-28)."""
    from scipy.interpolate import interp1d

    energy, let = _let_table(particle)
    return float(interp1d(energy, let)(E_MeV_u))


@lru_cache(maxsize=None)
def _radius_fit():
    rows = np.genfromtxt(DATA_DIR / "LET_b.dat", delimiter=",", dtype=float)
    let_kev_um = rows[:, 0] * 1e-3
    return np.poly1d(np.polyfit(np.log10(let_kev_um), rows[:, 1], 2))


def calc_track_radius_cm(LET_keV_um):
    """Rossomme track radius b [cm]: quadratic in log10(LET), floored at 20 um
    (stopping_power.py:31-42)."""
    b_cm = _radius_fit()(np.log10(LET_keV_um)) * 1e-3
    return max(b_cm, 2e-3)


def dose_rate_to_fluence_rate(dose_rate_Gy_s, E_MeV_u, particle="proton", air_density_kg_m3=None):
    """Dose rate in dry air [Gy/s] to fluence rate [cm^-2 s^-1] (stopping_power.py:45-56)."""
    let_kev_cm = E_MeV_u_to_LET_keV_um(E_MeV_u, particle) * 1e4
    density = AIR_DENSITY_KG_M3 if air_density_kg_m3 is None else air_density_kg_m3
    return dose_rate_Gy_s * JOULE_TO_KEV * (density * 1e-6) / let_kev_cm
